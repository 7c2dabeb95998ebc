/*
 * qmc_oracle.c -- CPU FP64 restatement of the reference's walker-evaluation
 * path (fabioaffinito/2degas, reptation.f).  TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP library, never linked into it.
 *
 * It follows the reference routines loop for loop, with the same pair order,
 * the same pivoting and the same inverse algorithm, so that agreement with it
 * means agreement with the reference's arithmetic up to rounding.  Each
 * function names the routine and lines it restates.
 */
#include "qmc_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

/* Arithmetic type of the evaluation.  The oracle is built twice: in double
 * (libqmc_oracle.so, the reference's own precision and loop order) and in
 * long double (libqmc_oracle_ld.so, -DORC_REAL="long double"): the same
 * sums with a 64-bit significand, which serve as the exact value of the
 * reference's formulas for the double inputs (tables, g-vectors,
 * positions) when the rounding error of an FP64 result is measured.  The
 * heap sort of the twist list and the random-number generator stay in
 * double and integers in both builds. */
#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
typedef ORC_REAL _Complex zc;

struct orc {
  int ntypes, np[2], ipf[2], ipl[2], nptot, npnorm;
  real hbs2m[2], el[2], eli[2], drt, drti, drti2;
  int ntab, ldut;
  real* ut;
  int32_t *ngrid, *iexp, *nk_ewald;
  real* tail;
  int iu2[2][2], iu3[2][2], iub[2][2], iv2[2][2];
  real v2value[2][2], v0;
  int nk;
  real* kvec;
  int irhok[2], nrhok, ntheta;
  real adrift, delta;
  int pp_dist[2][2];
  int ng;
  real* gvec; /* [ng][2] */
  qmc_info info;
  /* pair tables (distances, reptation.f:4127-4178) */
  int npairs;
  real *pp_r, *pp_byr, *pp_rvec, *pp_rem;
  int* pp_ind;
  real* rhok; /* [2][2*nk], column = type (reptation.f:4224-4243) */
  real var[2], vari[2];
};

#define UT(o, t, i, c) ((o)->ut[((size_t)(t)*4 + (c)) * (o)->ldut + (i)])

/* ------------------------------------------------------------------------ */
/* getf / getf_0: cubic spline on the uniform grid (reptation.f:4103-4125). */
static void getf(const orc* o, int tab, int ind, real rem, real* f,
                 real* df, real* ddf) {
  int t = tab - 1;
  int i = ind < o->ngrid[t] ? ind : o->ngrid[t];
  real rm = rem * o->drti;
  real t1 = UT(o, t, i, 0), t2 = UT(o, t, i, 1), t3 = UT(o, t, i, 2),
         t4 = UT(o, t, i, 3);
  *f = t1 + rm * (t2 + rm * (t3 + rm * t4));
  *df = (t2 + rm * (2 * t3 + rm * (3 * t4))) * o->drti;
  *ddf = (t3 + rm * (3 * t4)) * o->drti2;
}

static real ipow(real b, int e) {
  real r = 1.0;
  int n = e < 0 ? -e : e;
  for (int k = 0; k < n; k++) r *= b;
  return e < 0 ? 1.0 / r : r;
}

static real getf_0(const orc* o, int tab, int ind, real rem, real byr) {
  int t = tab - 1;
  int i = ind < o->ngrid[t] ? ind : o->ngrid[t];
  real rm = rem * o->drti;
  real f = UT(o, t, i, 0) +
             rm * (UT(o, t, i, 1) + rm * (UT(o, t, i, 2) + rm * UT(o, t, i, 3)));
  return f * ipow(byr, o->iexp[t]);
}

/* ------------------------------------------------------------------------ */
static int count_pairs(const orc* o, int it, int jt) {
  if (it == jt) return o->np[it] * (o->np[it] - 1) / 2;
  return o->np[it] * o->np[jt];
}

orc* orc_create(const qmc_tables* t, int* err) {
  *err = QMC_ERR_ARG;
  if (!t || t->ndim != 2 || t->ntypes < 1 || t->ntypes > 2) return NULL;
  orc* o = (orc*)calloc(1, sizeof(orc));
  o->ntypes = t->ntypes;
  int first = 0;
  o->npnorm = 0;
  for (int it = 0; it < o->ntypes; it++) {
    o->np[it] = t->np[it];
    o->hbs2m[it] = t->hbs2m[it];
    o->ipf[it] = first;
    o->ipl[it] = first + t->np[it] - 1;
    first += t->np[it];
    if (t->hbs2m[it] != 0.0) o->npnorm += t->np[it];
    else if (t->np[it] > 0) { free(o); return NULL; } /* static particles unsupported */
  }
  o->nptot = first;
  for (int d = 0; d < 2; d++) {
    o->el[d] = t->el[d];
    o->eli[d] = 1.0 / t->el[d];
  }
  o->drt = t->drt;
  o->drti = 1.0 / t->drt;
  o->drti2 = 2.0 * o->drti * o->drti;
  o->ntab = t->ntab;
  o->ldut = t->ldut;
  o->ut = (real*)malloc(sizeof(real) * 4 * (size_t)t->ldut * t->ntab);
  for (size_t k = 0; k < 4 * (size_t)t->ldut * t->ntab; k++) o->ut[k] = t->ut[k];
  o->ngrid = (int32_t*)malloc(sizeof(int32_t) * t->ntab);
  o->iexp = (int32_t*)malloc(sizeof(int32_t) * t->ntab);
  o->nk_ewald = (int32_t*)malloc(sizeof(int32_t) * t->ntab);
  o->tail = (real*)malloc(sizeof(real) * t->ntab);
  memcpy(o->ngrid, t->ngrid, sizeof(int32_t) * t->ntab);
  memcpy(o->iexp, t->iexp, sizeof(int32_t) * t->ntab);
  memcpy(o->nk_ewald, t->nk_ewald, sizeof(int32_t) * t->ntab);
  for (int k = 0; k < t->ntab; k++) o->tail[k] = t->tail[k];
  for (int it = 0; it < 2; it++)
    for (int jt = 0; jt < 2; jt++) {
      int k = it * 2 + jt;
      o->iu2[it][jt] = t->iu2[k];
      o->iu3[it][jt] = t->iu3[k];
      o->iub[it][jt] = t->iub[k];
      o->iv2[it][jt] = t->iv2[k];
      o->v2value[it][jt] = t->v2value[k];
      o->pp_dist[it][jt] = (t->iu2[k] || t->iu3[k] || t->iub[k] || t->iv2[k]);
    }
  o->v0 = t->v0;
  o->nk = t->nk;
  o->kvec = (real*)malloc(sizeof(real) * 2 * (t->nk > 0 ? t->nk : 1));
  for (int k = 0; k < 2 * t->nk; k++) o->kvec[k] = t->kvec[k];
  o->nrhok = t->nrhok;
  for (int it = 0; it < 2; it++) {
    o->irhok[it] = t->irhok[it];
    if (o->irhok[it] != 0 && o->irhok[it] != it + 1) { orc_destroy(o); return NULL; }
  }
  o->ntheta = t->ntheta;
  o->adrift = t->adrift;
  o->delta = t->delta;
  /* pair counters of the table routines must match distances' pair list */
  for (int it = 0; it < o->ntypes; it++)
    for (int jt = it; jt < o->ntypes; jt++) {
      if (count_pairs(o, it, jt) == 0 || !o->pp_dist[it][jt]) continue;
      for (int a = 0; a < o->ntypes; a++)
        for (int b = a; b < o->ntypes; b++) {
          if (count_pairs(o, a, b) == 0 || !o->pp_dist[a][b]) continue;
          if ((!o->iu2[it][jt]) != (!o->iu2[a][b]) || (!o->iu3[it][jt]) != (!o->iu3[a][b]) ||
              (!o->iub[it][jt]) != (!o->iub[a][b]) || (!o->iv2[it][jt]) != (!o->iv2[a][b])) {
            orc_destroy(o);
            return NULL;
          }
        }
    }
  int np2 = o->nptot * (o->nptot + 1) / 2 + 1;
  o->pp_r = (real*)calloc(np2, sizeof(real));
  o->pp_byr = (real*)calloc(np2, sizeof(real));
  o->pp_rem = (real*)calloc(np2, sizeof(real));
  o->pp_rvec = (real*)calloc(2 * np2, sizeof(real));
  o->pp_ind = (int*)calloc(np2, sizeof(int));
  o->rhok = (real*)calloc(2 * 2 * (size_t)(o->nk > 0 ? o->nk : 1), sizeof(real));

  /* property slots (reptation.f:577-635) */
  qmc_info* in = &o->info;
  memset(in, 0, sizeof(*in));
  in->nptot = o->nptot;
  in->npnorm = o->npnorm;
  in->jetot = 1; in->jltf = 2; in->jacc = 3; in->jpot0 = 4; in->jkin0 = 5;
  in->je2 = 6; in->jgg = 7;
  int i = 7;
  for (int it = 0; it < o->ntypes; it++) if (o->hbs2m[it] != 0) in->jpot[it] = ++i;
  for (int it = 0; it < o->ntypes; it++) if (o->hbs2m[it] != 0) in->jkin[it] = ++i;
  for (int it = 0; it < o->ntypes; it++) if (o->hbs2m[it] != 0) in->jun[it] = ++i;
  in->jefp = ++i;
  in->n_scalar_props = i;
  in->jrhok = i + 1;
  in->n_props = i + 2 * o->nk * o->nrhok;
  int has_det = 0;
  for (int it = 0; it < o->ntypes; it++)
    for (int jt = it; jt < o->ntypes; jt++) has_det += o->iub[it][jt] != 0;
  in->has_det = has_det != 0;
  *err = QMC_OK;
  return o;
}

void orc_destroy(orc* o) {
  if (!o) return;
  free(o->ut); free(o->ngrid); free(o->iexp); free(o->nk_ewald); free(o->tail);
  free(o->kvec); free(o->gvec);
  free(o->pp_r); free(o->pp_byr); free(o->pp_rem); free(o->pp_rvec); free(o->pp_ind);
  free(o->rhok);
  free(o);
}

int orc_info(const orc* o, qmc_info* out) {
  *out = o->info;
  out->ng = o->ng;
  return QMC_OK;
}

int orc_set_twist(orc* o, const double* gvec, int ng) {
  free(o->gvec);
  o->gvec = (real*)malloc(sizeof(real) * 2 * (ng > 0 ? ng : 1));
  for (int k = 0; k < 2 * ng; k++) o->gvec[k] = gvec[k];
  o->ng = ng;
  return QMC_OK;
}

/* ------------------------------------------------------------------------ */
/* SORTTF: index heap sort by ascending A (reptation.f:9514-9557), 1-based. */
static void sorttf(int n1, const double* A, int* INDEX) {
  int N = n1, I1, I2, I3, I22 = 0, I33, I222;
  double AI;
  for (I1 = 2; I1 <= N; I1++) {
    I3 = I1;
    I33 = INDEX[I3];
    AI = A[I33];
    for (;;) {
      I2 = I3 / 2;
      if (I2 <= 0) break;
      I22 = INDEX[I2];
      if (AI <= A[I22]) break;
      INDEX[I3] = I22;
      I3 = I2;
    }
    INDEX[I3] = I33;
  }
  for (;;) {
    I3 = INDEX[N];
    INDEX[N] = INDEX[1];
    AI = A[I3];
    N = N - 1;
    if (N - 1 <= 0) break;
    I1 = 1;
    for (;;) {
      I2 = I1 + I1;
      if (I2 <= N) I22 = INDEX[I2];
      if (I2 - N < 0) {
        I222 = INDEX[I2 + 1];
        if (A[I22] - A[I222] < 0) {
          I2 = I2 + 1;
          I22 = I222;
        }
      } else if (I2 - N > 0) {
        INDEX[I1] = I3;
        goto next;
      }
      if (AI - A[I22] < 0) {
        INDEX[I1] = I22;
        I1 = I2;
      } else {
        INDEX[I1] = I3;
        goto next;
      }
    }
  next:;
  }
  INDEX[1] = I3;
}

/* twistg (reptation.f:9818-9861) when variant != 0, tabc's association
 * theta = ith*(2*pi*eli/ntheta) (reptation.f:9578-9583) when variant == 0;
 * ntheta == 0 gives input's untwisted list (reptation.f:674-699). */
static int twist_gvec(const double* kvec, int nk, const double* el, int ntheta,
                      int ith, int jth, double* gvec_out, int variant) {
  int n = 2 * nk + 1;
  double pi = acos(-1.0);
  if (ntheta == 0) {
    gvec_out[0] = 0.0; gvec_out[1] = 0.0;
    for (int i = 0; i < nk; i++) {
      gvec_out[2 * (2 * i + 1) + 0] = kvec[2 * i + 0];
      gvec_out[2 * (2 * i + 1) + 1] = kvec[2 * i + 1];
      gvec_out[2 * (2 * i + 2) + 0] = -kvec[2 * i + 0];
      gvec_out[2 * (2 * i + 2) + 1] = -kvec[2 * i + 1];
    }
    return QMC_OK;
  }
  double theta[2];
  int ij[2] = {ith, jth};
  for (int d = 0; d < 2; d++) {
    double eli = 1.0 / el[d];
    if (variant) theta[d] = ij[d] * 2.0 * pi * eli / ntheta;
    else theta[d] = ij[d] * (2.0 * pi * eli / ntheta);
  }
  double* gv = (double*)malloc(sizeof(double) * 2 * (n + 1));
  double* gn = (double*)malloc(sizeof(double) * (n + 1));
  int* idx = (int*)malloc(sizeof(int) * (n + 1));
  /* 1-based like the reference */
  gn[1] = 0.0;
  for (int d = 0; d < 2; d++) {
    gv[2 * 1 + d] = theta[d];
    gn[1] = gn[1] + gv[2 * 1 + d] * gv[2 * 1 + d];
  }
  idx[1] = 1;
  for (int k = nk; k >= 1; k--) {
    gn[2 * k + 1] = 0.0;
    gn[2 * k] = 0.0;
    for (int d = 0; d < 2; d++) {
      gv[2 * (2 * k + 1) + d] = -kvec[2 * (k - 1) + d] + theta[d];
      gv[2 * (2 * k) + d] = kvec[2 * (k - 1) + d] + theta[d];
      gn[2 * k + 1] = gn[2 * k + 1] + gv[2 * (2 * k + 1) + d] * gv[2 * (2 * k + 1) + d];
      gn[2 * k] = gn[2 * k] + gv[2 * (2 * k) + d] * gv[2 * (2 * k) + d];
    }
    idx[2 * k] = 2 * k;
    idx[2 * k + 1] = 2 * k + 1;
  }
  sorttf(n, gn, idx);
  for (int k = 1; k <= n; k++)
    for (int d = 0; d < 2; d++) gvec_out[2 * (k - 1) + d] = gv[2 * idx[k] + d];
  free(gv); free(gn); free(idx);
  return QMC_OK;
}

int orc_twist_gvec(const double* kvec, int nk, const double* el, int ntheta,
                   int ith, int jth, double* gvec_out, int variant) {
  return twist_gvec(kvec, nk, el, ntheta, ith, jth, gvec_out, variant);
}

/* ------------------------------------------------------------------------ */
/* distances (reptation.f:4127-4178) and rhok (reptation.f:4224-4243).      */
static void distances(orc* o, const real* x) {
  int ij = 0;
  for (int it = 0; it < o->ntypes; it++) {
    for (int jt = it; jt < o->ntypes; jt++) {
      if (!o->pp_dist[it][jt]) continue;
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++) {
        int j0 = (jt == it) ? ip + 1 : o->ipf[jt];
        for (int jp = j0; jp <= o->ipl[jt]; jp++) {
          real r2 = 0.0;
          for (int d = 0; d < 2; d++) {
            real v = x[2 * ip + d] - x[2 * jp + d];
            v = v - o->el[d] * round(o->eli[d] * v);
            o->pp_rvec[2 * ij + d] = v;
            r2 = r2 + v * v;
          }
          real r = sqrt(r2);
          o->pp_r[ij] = r;
          o->pp_ind[ij] = (int)(r * o->drti);
          o->pp_rem[ij] = r - o->pp_ind[ij] * o->drt;
          o->pp_byr[ij] = 1.0 / r;
          ij++;
        }
      }
    }
  }
  o->npairs = ij;
  if (o->nrhok != 0) {
    for (int it = 0; it < o->ntypes; it++) {
      if (!o->irhok[it]) continue;
      real* rk = o->rhok + (size_t)it * 2 * o->nk;
      for (int k = 0; k < 2 * o->nk; k++) rk[k] = 0.0;
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++)
        for (int ik = 0; ik < o->nk; ik++) {
          real kr = 0.0;
          for (int d = 0; d < 2; d++) kr = kr + o->kvec[2 * ik + d] * x[2 * ip + d];
          rk[2 * ik] += cos(kr);
          rk[2 * ik + 1] += sin(kr);
        }
    }
  }
}

/* Walk the pair list the way every table routine does: same-type pairs
 * i<j of type it, then cross pairs (it,jt>it), each block only when the
 * routine's table for it is declared (reptation.f:2993-3038). */
#define FOR_PAIRS(o, TAB, ...)                                             \
  do {                                                                     \
    int ij_ = 0;                                                           \
    for (int it = 0; it < (o)->ntypes; it++) {                             \
      for (int jt = it; jt < (o)->ntypes; jt++) {                          \
        int tab = (o)->TAB[it][jt];                                        \
        if (!tab) continue;                                                \
        for (int ip = (o)->ipf[it]; ip <= (o)->ipl[it]; ip++) {            \
          int j0_ = (jt == it) ? ip + 1 : (o)->ipf[jt];                    \
          for (int jp = j0_; jp <= (o)->ipl[jt]; jp++) {                   \
            int ij = ij_++;                                                \
            __VA_ARGS__                                                    \
          }                                                                \
        }                                                                  \
      }                                                                    \
    }                                                                      \
  } while (0)

typedef struct wf {
  real ltf;
  real* g; /* [nptot][2] */
  real h[2];
  real grad2_ph;
  int guard;
} wf;

/* two_body (reptation.f:2981-3046) */
static void two_body(orc* o, wf* w) {
  real ltf = 0.0, h[2] = {0, 0};
  real* g = (real*)calloc(2 * (size_t)o->nptot, sizeof(real));
  FOR_PAIRS(o, iu2, {
    real t, dt, ddt;
    getf(o, tab, o->pp_ind[ij], o->pp_rem[ij], &t, &dt, &ddt);
    ltf = ltf + t;
    dt = dt * o->pp_byr[ij];
    for (int d = 0; d < 2; d++) {
      g[2 * ip + d] += dt * o->pp_rvec[2 * ij + d];
      g[2 * jp + d] -= dt * o->pp_rvec[2 * ij + d];
    }
    if (it == jt) {
      h[it] += 2 * (ddt + (2 - 1) * dt);
    } else {
      h[it] += ddt + (2 - 1) * dt;
      h[jt] += ddt + (2 - 1) * dt;
    }
  });
  w->ltf += ltf;
  for (int i = 0; i < 2 * o->nptot; i++) w->g[i] += g[i];
  for (int it = 0; it < o->ntypes; it++) w->h[it] += h[it];
  free(g);
}

/* three_body + makeg + dotg (reptation.f:2777-2979) */
static void three_body(orc* o, wf* w) {
  int any = 0;
  for (int it = 0; it < o->ntypes; it++)
    for (int jt = it; jt < o->ntypes; jt++) any += o->iu3[it][jt] != 0;
  if (!any) return;
  int n = o->nptot;
  real* G = (real*)calloc(2 * (size_t)n, sizeof(real));
  real* dg = (real*)calloc(4 * (size_t)n * n, sizeof(real));
  real* ddg = (real*)calloc(2 * (size_t)n * n, sizeof(real));
#define DG(jd, id, jp_, ip_) dg[((((size_t)(ip_)) * n + (jp_)) * 2 + (id)) * 2 + (jd)]
#define DDG(id, jp_, ip_) ddg[(((size_t)(ip_)) * n + (jp_)) * 2 + (id)]
  FOR_PAIRS(o, iu3, {
    real t, dt, ddt;
    getf(o, tab, o->pp_ind[ij], o->pp_rem[ij], &t, &dt, &ddt);
    const real* delx = o->pp_rvec + 2 * ij;
    real rij = o->pp_byr[ij];
    /* makeg (reptation.f:2884-2926) */
    {
      real aux1 = t, aux2, aux3;
      for (int d = 0; d < 2; d++) {
        aux2 = -aux1 * delx[d];
        G[2 * ip + d] += aux2;
        G[2 * jp + d] -= aux2;
      }
      aux2 = dt * rij;
      for (int id = 0; id < 2; id++) {
        aux3 = aux1 + aux2 * delx[id] * delx[id];
        DG(id, id, jp, ip) = -aux3;
        DG(id, id, ip, ip) += aux3;
        DG(id, id, jp, jp) += aux3;
        for (int jd = id + 1; jd < 2; jd++) {
          aux3 = aux2 * delx[id] * delx[jd];
          DG(jd, id, jp, ip) = -aux3;
          DG(jd, id, ip, ip) += aux3;
          DG(jd, id, jp, jp) += aux3;
        }
      }
      aux2 = (2 + 1) * dt * rij + ddt;
      for (int id = 0; id < 2; id++) {
        aux3 = aux2 * delx[id];
        DDG(id, jp, ip) = aux3;
        DDG(id, ip, ip) -= aux3;
        DDG(id, jp, jp) += aux3;
      }
    }
    /* remove a two-body term (reptation.f:2822-2832) */
    real r = o->pp_r[ij];
    ddt = 2 * (t * t + 4 * r * t * dt + r * r * (dt * dt + ddt * t));
    dt = 2 * t * (t + r * dt);
    t = (t * r) * (t * r);
    w->ltf += t;
    for (int d = 0; d < 2; d++) {
      w->g[2 * ip + d] += dt * delx[d];
      w->g[2 * jp + d] -= dt * delx[d];
    }
    if (it == jt) {
      w->h[it] += 2 * (ddt + (2 - 1) * dt);
    } else {
      w->h[it] += ddt + (2 - 1) * dt;
      w->h[jt] += ddt + (2 - 1) * dt;
    }
  });
  /* dotg (reptation.f:2928-2979) */
  for (int ip = 0; ip < n; ip++)
    for (int d = 0; d < 2; d++) w->ltf -= 0.5 * G[2 * ip + d] * G[2 * ip + d];
  real* gn = w->g;
  for (int it = 0; it < o->ntypes; it++) {
    if (o->iu3[it][it]) {
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++) {
        for (int id = 0; id < 2; id++) {
          gn[2 * ip + id] += G[2 * ip + id] * DG(id, id, ip, ip);
          w->h[it] -= DG(id, id, ip, ip) * DG(id, id, ip, ip);
          w->h[it] -= G[2 * ip + id] * DDG(id, ip, ip);
          for (int jd = id + 1; jd < 2; jd++) {
            gn[2 * ip + id] += G[2 * ip + jd] * DG(jd, id, ip, ip);
            gn[2 * ip + jd] += G[2 * ip + id] * DG(jd, id, ip, ip);
            w->h[it] -= 2 * DG(jd, id, ip, ip) * DG(jd, id, ip, ip);
          }
        }
        for (int jp = ip + 1; jp <= o->ipl[it]; jp++) {
          for (int id = 0; id < 2; id++) {
            gn[2 * ip + id] += G[2 * jp + id] * DG(id, id, jp, ip);
            gn[2 * jp + id] += G[2 * ip + id] * DG(id, id, jp, ip);
            w->h[it] -= 2 * DG(id, id, jp, ip) * DG(id, id, jp, ip);
            w->h[it] += G[2 * ip + id] * DDG(id, jp, ip);
            w->h[it] -= G[2 * jp + id] * DDG(id, jp, ip);
            for (int jd = id + 1; jd < 2; jd++) {
              gn[2 * ip + id] += G[2 * jp + jd] * DG(jd, id, jp, ip);
              gn[2 * ip + jd] += G[2 * jp + id] * DG(jd, id, jp, ip);
              gn[2 * jp + id] += G[2 * ip + jd] * DG(jd, id, jp, ip);
              gn[2 * jp + jd] += G[2 * ip + id] * DG(jd, id, jp, ip);
              w->h[it] -= 4 * DG(jd, id, jp, ip) * DG(jd, id, jp, ip);
            }
          }
        }
      }
    }
    for (int jt = it + 1; jt < o->ntypes; jt++) {
      if (!o->iu3[it][jt]) continue;
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++)
        for (int jp = o->ipf[jt]; jp <= o->ipl[jt]; jp++)
          for (int id = 0; id < 2; id++) {
            gn[2 * ip + id] += G[2 * jp + id] * DG(id, id, jp, ip);
            gn[2 * jp + id] += G[2 * ip + id] * DG(id, id, jp, ip);
            w->h[it] -= DG(id, id, jp, ip) * DG(id, id, jp, ip);
            w->h[jt] -= DG(id, id, jp, ip) * DG(id, id, jp, ip);
            w->h[it] -= G[2 * jp + id] * DDG(id, jp, ip);
            w->h[jt] += G[2 * ip + id] * DDG(id, jp, ip);
            for (int jd = id + 1; jd < 2; jd++) {
              gn[2 * ip + id] += G[2 * jp + jd] * DG(jd, id, jp, ip);
              gn[2 * ip + jd] += G[2 * jp + id] * DG(jd, id, jp, ip);
              gn[2 * jp + id] += G[2 * ip + jd] * DG(jd, id, jp, ip);
              gn[2 * jp + jd] += G[2 * ip + id] * DG(jd, id, jp, ip);
              w->h[it] -= 2 * DG(jd, id, jp, ip) * DG(jd, id, jp, ip);
              w->h[jt] -= 2 * DG(jd, id, jp, ip) * DG(jd, id, jp, ip);
            }
          }
    }
  }
#undef DG
#undef DDG
  free(G); free(dg); free(ddg);
}

/* ------------------------------------------------------------------------ */
/* LINPACK zgefa / zgedi(job=11) restated (reptation.f:9227-9472);
 * column-major a(row, col) = a[col*n + row]. */
static real cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); }

static int zgefa(zc* a, int n, int* ipvt) {
  int info = 0;
#define A(r, c) a[(size_t)(c) * n + (r)]
  for (int k = 0; k < n - 1; k++) {
    int l = k;
    real smax = cabs1(A(k, k));
    for (int i = k + 1; i < n; i++)
      if (cabs1(A(i, k)) > smax) { l = i; smax = cabs1(A(i, k)); }
    ipvt[k] = l;
    if (cabs1(A(l, k)) == 0.0) { info = k + 1; continue; }
    if (l != k) { zc t = A(l, k); A(l, k) = A(k, k); A(k, k) = t; }
    zc t = -1.0 / A(k, k);
    for (int i = k + 1; i < n; i++) A(i, k) = t * A(i, k);
    for (int j = k + 1; j < n; j++) {
      t = A(l, j);
      if (l != k) { A(l, j) = A(k, j); A(k, j) = t; }
      if (cabs1(t) != 0.0)
        for (int i = k + 1; i < n; i++) A(i, j) = A(i, j) + t * A(i, k);
    }
  }
  if (n >= 1) {
    ipvt[n - 1] = n - 1;
    if (cabs1(A(n - 1, n - 1)) == 0.0) info = n;
  }
  return info;
}

static void zgedi(zc* a, int n, const int* ipvt, zc det[2], zc* work) {
  det[0] = 1.0;
  det[1] = 0.0;
  for (int i = 0; i < n; i++) {
    if (ipvt[i] != i) det[0] = -det[0];
    det[0] = A(i, i) * det[0];
    if (cabs1(det[0]) == 0.0) break;
    while (cabs1(det[0]) < 1.0) { det[0] = 10.0 * det[0]; det[1] = det[1] - 1.0; }
    while (cabs1(det[0]) >= 10.0) { det[0] = det[0] / 10.0; det[1] = det[1] + 1.0; }
  }
  for (int k = 0; k < n; k++) {
    A(k, k) = 1.0 / A(k, k);
    zc t = -A(k, k);
    for (int i = 0; i < k; i++) A(i, k) = t * A(i, k);
    for (int j = k + 1; j < n; j++) {
      t = A(k, j);
      A(k, j) = 0.0;
      if (cabs1(t) != 0.0)
        for (int i = 0; i <= k; i++) A(i, j) = A(i, j) + t * A(i, k);
    }
  }
  for (int k = n - 2; k >= 0; k--) {
    for (int i = k + 1; i < n; i++) { work[i] = A(i, k); A(i, k) = 0.0; }
    for (int j = k + 1; j < n; j++) {
      zc t = work[j];
      if (cabs1(t) != 0.0)
        for (int i = 0; i < n; i++) A(i, k) = A(i, k) + t * A(i, j);
    }
    int l = ipvt[k];
    if (l != k)
      for (int i = 0; i < n; i++) { zc t = A(i, k); A(i, k) = A(i, l); A(i, l) = t; }
  }
#undef A
}

/* slater_excitations + paho_orbitals (reptation.f:9967-10363) */
static int slater(orc* o, const real* x, wf* w) {
  if (!o->info.has_det) return QMC_OK;
  int n = o->nptot;
  for (int it = 0; it < o->ntypes; it++)
    if (o->np[it] > o->ng) return QMC_ERR_NOTWIST;
  real* qx = (real*)malloc(sizeof(real) * 2 * n);
  real* qa = (real*)calloc(4 * (size_t)n * n, sizeof(real));
  real* qb = (real*)calloc(2 * (size_t)n * n, sizeof(real));
#define QA(jd, id, jp_, ip_) qa[((((size_t)(ip_)) * n + (jp_)) * 2 + (id)) * 2 + (jd)]
#define QB(id, jp_, ip_) qb[(((size_t)(ip_)) * n + (jp_)) * 2 + (id)]
  for (int ip = 0; ip < n; ip++)
    for (int d = 0; d < 2; d++) {
      qx[2 * ip + d] = x[2 * ip + d];
      QA(d, d, ip, ip) = 1.0;
    }
  FOR_PAIRS(o, iub, {
    real t, dt, ddt;
    getf(o, tab, o->pp_ind[ij], o->pp_rem[ij], &t, &dt, &ddt);
    for (int d = 0; d < 2; d++) {
      real v = t * o->pp_rvec[2 * ij + d];
      qx[2 * ip + d] += v;
      qx[2 * jp + d] -= v;
    }
  });
  /* per-spin orbital matrices, inverse and derivative matrices */
  zc* inv[2] = {NULL, NULL};
  zc* dA[2] = {NULL, NULL};  /* [2][npa][npa] : d/dq_b A(k, i)   */
  zc* ddA[2] = {NULL, NULL}; /* [2][2][npa][npa]                 */
  for (int jt = 0; jt < o->ntypes; jt++) {
    int m = o->np[jt];
    int mm = m > 0 ? m : 1;
    inv[jt] = (zc*)malloc(sizeof(zc) * mm * mm);
    dA[jt] = (zc*)malloc(sizeof(zc) * 2 * mm * mm);
    ddA[jt] = (zc*)malloc(sizeof(zc) * 4 * mm * mm);
    for (int jk = 0; jk < m; jk++) {
      const real* gk = o->gvec + 2 * jk;
      for (int jp = o->ipf[jt]; jp <= o->ipl[jt]; jp++) {
        real kr = 0.0;
        for (int d = 0; d < 2; d++) kr = kr + qx[2 * jp + d] * gk[d];
        int i = jp - o->ipf[jt];
        zc z = cos(kr) + I * sin(kr);
        inv[jt][(size_t)i * m + jk] = z;
        for (int d = 0; d < 2; d++) dA[jt][((size_t)d * m + i) * m + jk] = (I * gk[d]) * z;
        for (int id = 0; id < 2; id++)
          for (int jd = 0; jd < 2; jd++)
            ddA[jt][(((size_t)(id * 2 + jd)) * m + i) * m + jk] = -(gk[id] * gk[jd]) * z;
      }
    }
    int* ipvt = (int*)malloc(sizeof(int) * mm);
    zc* wk = (zc*)malloc(sizeof(zc) * mm);
    int info = zgefa(inv[jt], m, ipvt);
    if (info != 0) {
      free(ipvt); free(wk);
      for (int k = 0; k <= jt; k++) { free(inv[k]); free(dA[k]); free(ddA[k]); }
      free(qx); free(qa); free(qb);
      return QMC_ERR_SINGULAR;
    }
    zc det[2];
    zgedi(inv[jt], m, ipvt, det, wk);
    w->ltf -= creal(log(det[0]) + det[1] * log((real)10));
    free(ipvt); free(wk);
  }
  real ww = exp(-w->ltf * 2);
  if (ww <= 0.0) {
    w->guard = 1;
    for (int k = 0; k < o->ntypes; k++) { free(inv[k]); free(dA[k]); free(ddA[k]); }
    free(qx); free(qa); free(qb);
    return QMC_OK;
  }
  FOR_PAIRS(o, iub, {
    real t, dt, ddt;
    getf(o, tab, o->pp_ind[ij], o->pp_rem[ij], &t, &dt, &ddt);
    const real* rv = o->pp_rvec + 2 * ij;
    real byr = o->pp_byr[ij];
    for (int id = 0; id < 2; id++) {
      real w2 = (dt * byr * (2 + 1) + ddt) * rv[id];
      QB(id, ip, ip) += w2;
      QB(id, jp, jp) -= w2;
      QB(id, jp, ip) += w2;
      QB(id, ip, jp) -= w2;
      QA(id, id, ip, ip) += t;
      QA(id, id, jp, jp) += t;
      QA(id, id, jp, ip) -= t;
      QA(id, id, ip, jp) -= t;
      for (int jd = 0; jd < 2; jd++) {
        real w3 = dt * byr * rv[jd] * rv[id];
        QA(jd, id, ip, ip) += w3;
        QA(jd, id, jp, jp) += w3;
        QA(jd, id, jp, ip) -= w3;
        QA(jd, id, ip, jp) -= w3;
      }
    }
  });
  real aux = 0.0;
  for (int jt = 0; jt < o->ntypes; jt++) {
    int m = o->np[jt];
    if (m == 0) continue;
    const zc* Ai = inv[jt]; /* Ai(l, k) = Ai[k*m + l]: rows particles, cols orbitals */
    /* zv(b, l, j) = sum_k Ai(l,k) dA(b,k,j) (reptation.f:10182-10194) */
    zc* zv = (zc*)calloc(2 * (size_t)m * m, sizeof(zc));
#define ZV(b, l, j) zv[((size_t)(b) * m + (j)) * m + (l)]
    for (int jf = 0; jf < m; jf++)
      for (int lf = 0; lf < m; lf++)
        for (int b = 0; b < 2; b++) {
          zc s = 0.0;
          for (int kf = 0; kf < m; kf++)
            s = s + Ai[(size_t)kf * m + lf] * dA[jt][((size_t)b * m + jf) * m + kf];
          ZV(b, lf, jf) = s;
        }
    for (int it = 0; it < o->ntypes; it++) {
      if (!o->iub[it][jt]) continue;
      real aux2 = 0.0;
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++)
        for (int id = 0; id < 2; id++) {
          zc z = 0.0;
          for (int jp = o->ipf[jt]; jp <= o->ipl[jt]; jp++) {
            int jf = jp - o->ipf[jt];
            for (int jd = 0; jd < 2; jd++) z = z + ZV(jd, jf, jf) * QA(jd, id, jp, ip);
          }
          w->g[2 * ip + id] -= creal(z);
          aux2 = aux2 + cimag(z) * cimag(z);
        }
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++)
        for (int jp = o->ipf[jt]; jp <= o->ipl[jt]; jp++) {
          int jf = jp - o->ipf[jt];
          for (int jd = 0; jd < 2; jd++) w->h[it] -= creal(ZV(jd, jf, jf)) * QB(jd, ip, jp);
        }
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++)
        for (int jp = o->ipf[jt]; jp <= o->ipl[jt]; jp++) {
          int jf = jp - o->ipf[jt];
          for (int id = 0; id < 2; id++)
            for (int jd = 0; jd < 2; jd++) {
              zc z1 = 0.0;
              for (int kf = 0; kf < m; kf++)
                z1 = z1 + Ai[(size_t)kf * m + jf] *
                              ddA[jt][(((size_t)(id * 2 + jd)) * m + jf) * m + kf];
              real z2 = 0.0;
              for (int kd = 0; kd < 2; kd++) z2 = z2 + QA(kd, jd, jp, ip) * QA(kd, id, jp, ip);
              w->h[it] -= creal(z1) * z2;
            }
        }
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++)
        for (int jp = o->ipf[jt]; jp <= o->ipl[jt]; jp++) {
          int jf = jp - o->ipf[jt];
          for (int kp = o->ipf[jt]; kp <= o->ipl[jt]; kp++) {
            int kf = kp - o->ipf[jt];
            for (int id = 0; id < 2; id++) {
              zc z1 = 0.0, z2 = 0.0;
              for (int jd = 0; jd < 2; jd++) z1 = z1 + QA(jd, id, jp, ip) * ZV(jd, kf, jf);
              for (int jd = 0; jd < 2; jd++) z2 = z2 + QA(jd, id, kp, ip) * ZV(jd, jf, kf);
              w->h[it] += creal(z1 * z2);
            }
          }
        }
      aux = aux + aux2 * o->hbs2m[it];
    }
#undef ZV
    free(zv);
  }
  w->grad2_ph = aux;
  for (int k = 0; k < o->ntypes; k++) { free(inv[k]); free(dA[k]); free(ddA[k]); }
  free(qx); free(qa); free(qb);
#undef QA
#undef QB
  return QMC_OK;
}

/* ------------------------------------------------------------------------ */
/* kinetic (reptation.f:2729-2775) */
static void kinetic(orc* o, wf* w, real* p) {
  const qmc_info* in = &o->info;
  int n_jgg = 0;
  p[in->jkin0 - 1] = 0.0;
  p[in->jgg - 1] = 0.0;
  for (int it = 0; it < o->ntypes; it++) {
    if (o->hbs2m[it] == 0.0) continue;
    real g2 = 0.0;
    if (o->adrift == 0.0) {
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++)
        for (int d = 0; d < 2; d++) g2 = g2 + w->g[2 * ip + d] * w->g[2 * ip + d];
      p[in->jgg - 1] = 1.0;
      n_jgg = 1;
    } else {
      real vari = 1.0 / (2.0 * o->hbs2m[it] * o->delta);
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++) {
        real x2 = 0.0;
        for (int d = 0; d < 2; d++) x2 = x2 + w->g[2 * ip + d] * w->g[2 * ip + d];
        real gg = x2 * o->adrift * o->delta * o->delta * vari;
        gg = (-1.0 + sqrt(1.0 + 2.0 * gg)) / gg;
        p[in->jgg - 1] += gg;
        n_jgg++;
        for (int d = 0; d < 2; d++) w->g[2 * ip + d] *= gg;
        g2 = g2 + x2;
      }
    }
    real kin = o->hbs2m[it] * (w->h[it] - g2);
    p[in->jkin0 - 1] += kin;
    p[in->jkin[it] - 1] = kin / o->np[it];
    p[in->jun[it] - 1] = 1.0 / sqrt(g2);
  }
  p[in->jkin0 - 1] = p[in->jkin0 - 1] / o->npnorm;
  p[in->jgg - 1] = p[in->jgg - 1] / n_jgg;
  if (o->ntheta != 0) p[in->jefp - 1] = w->grad2_ph / o->npnorm;
  else p[in->jefp - 1] = 0.0;
}

/* potential (reptation.f:3986-4101), egas subset: pair potential with tails,
 * k-space Ewald through rhok, v0. */
static void potential(orc* o, real* p) {
  const qmc_info* in = &o->info;
  real p2[3] = {0, 0, 0}; /* [0]=total, [1+it] per type */
  int ij = 0;
  for (int it = 0; it < o->ntypes; it++) {
    for (int jt = it; jt < o->ntypes; jt++) {
      int tab = o->iv2[it][jt];
      if (!tab) continue;
      real v = o->v2value[it][jt];
      if (jt == it) {
        real aux = o->np[it] * (o->np[it] - 1) * 0.5 * o->tail[tab - 1] * v;
        p2[1 + it] += aux;
        p2[0] += aux;
      } else {
        real aux = o->np[it] * o->np[jt] * o->tail[tab - 1] * v;
        p2[1 + it] += 0.5 * aux;
        p2[1 + jt] += 0.5 * aux;
        p2[0] += aux;
      }
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++) {
        int j0 = (jt == it) ? ip + 1 : o->ipf[jt];
        for (int jp = j0; jp <= o->ipl[jt]; jp++) {
          real f = getf_0(o, tab, o->pp_ind[ij], o->pp_rem[ij], o->pp_byr[ij]);
          if (jt == it) {
            p2[1 + it] += f * v;
            p2[0] += f * v;
          } else {
            p2[1 + it] += 0.5 * f * v;
            p2[1 + jt] += 0.5 * f * v;
            p2[0] += f * v;
          }
          ij++;
        }
      }
    }
  }
  for (int it = 0; it < o->ntypes; it++)
    for (int jt = 0; jt < o->ntypes; jt++) {
      int tab = o->iv2[it][jt];
      if (!tab || o->nk_ewald[tab - 1] == 0) continue;
      int t = tab - 1, ng = o->ngrid[t], nke = o->nk_ewald[t];
      const real* ri = o->rhok + (size_t)it * 2 * o->nk;
      const real* rj = o->rhok + (size_t)jt * 2 * o->nk;
      real aux = 0.0;
      for (int ik = 0; ik < nke; ik++)
        aux = aux + UT(o, t, ik + 1 + ng, 0) *
                        (ri[2 * ik] * rj[2 * ik] + ri[2 * ik + 1] * rj[2 * ik + 1]);
      real vk0 = UT(o, t, ng + nke + 1, 0);
      p2[1 + it] += aux + vk0 * o->npnorm;
      p2[0] += aux + vk0 * o->npnorm;
    }
  for (int it = 0; it < o->ntypes; it++)
    if (in->jpot[it] != 0) p[in->jpot[it] - 1] = p2[1 + it] / o->np[it];
  p[in->jpot0 - 1] = (p2[0] + o->v0) / o->npnorm;
}

/* compute_properties(1) (reptation.f:2661-2727) */
static int eval_real(orc* o, const real* x, real* props, real* g, real* h,
                     real* grad2_ph, int32_t* status) {
  const qmc_info* in = &o->info;
  distances(o, x);
  wf w;
  memset(&w, 0, sizeof(w));
  w.g = g;
  for (int i = 0; i < 2 * o->nptot; i++) g[i] = 0.0;
  w.grad2_ph = grad2_ph ? *grad2_ph : 0.0;
  /* trial_function (reptation.f:3048-3063) */
  two_body(o, &w);
  three_body(o, &w);
  int rc = slater(o, x, &w);
  if (rc != QMC_OK) return rc;
  real jacc_saved = props[in->jacc - 1];
  real* p = props;
  p[in->jltf - 1] = w.ltf;
  for (int it = 0; it < o->ntypes; it++) h[it] = w.h[it];
  kinetic(o, &w, p);
  for (int it = 0; it < o->ntypes; it++) h[it] = w.h[it];
  potential(o, p);
  p[in->jetot - 1] = p[in->jpot0 - 1] + p[in->jkin0 - 1] + p[in->jefp - 1];
  p[in->je2 - 1] = p[in->jetot - 1] * p[in->jetot - 1];
  p[in->jacc - 1] = jacc_saved;
  int i = in->jrhok - 1;
  for (int it = 0; it < o->ntypes; it++) {
    if (!o->irhok[it]) continue;
    const real* rk = o->rhok + (size_t)it * 2 * o->nk;
    for (int ik = 0; ik < o->nk; ik++) {
      p[i] = rk[2 * ik];
      p[i + 1] = rk[2 * ik + 1];
      i += 2;
    }
  }
  if (grad2_ph) *grad2_ph = w.grad2_ph;
  if (status) *status = w.guard;
  return QMC_OK;
}

int orc_eval(orc* o, const double* x, double* props, double* g, double* h,
             double* grad2_ph, int32_t* status) {
  const int n2 = 2 * o->nptot, np_ = o->info.n_props;
  real* xr = (real*)malloc(sizeof(real) * (n2 + np_ + n2 + 2));
  real *pr = xr + n2, *gr = pr + np_, *hr = gr + n2, g2r = grad2_ph ? *grad2_ph : 0.0;
  for (int i = 0; i < n2; i++) xr[i] = x[i];
  for (int i = 0; i < np_; i++) pr[i] = props[i];
  hr[0] = h[0]; hr[1] = h[1];
  int rc = eval_real(o, xr, pr, gr, hr, grad2_ph ? &g2r : NULL, status);
  if (rc == QMC_OK) {
    for (int i = 0; i < np_; i++) props[i] = (double)pr[i];
    for (int i = 0; i < n2; i++) g[i] = (double)gr[i];
    for (int it = 0; it < o->ntypes; it++) h[it] = (double)hr[it];
    if (grad2_ph) *grad2_ph = (double)g2r;
  }
  free(xr);
  return rc;
}

int orc_eval_batch(orc* o, int nw, const double* x, double* props, double* g,
                   double* h, double* grad2_ph, int32_t* status) {
  const qmc_info* in = &o->info;
  for (int w = 0; w < nw; w++) {
    int rc = orc_eval(o, x + (size_t)w * 2 * o->nptot, props + (size_t)w * in->n_props,
                      g + (size_t)w * 2 * o->nptot, h + (size_t)w * 2,
                      grad2_ph ? grad2_ph + w : NULL, status ? status + w : NULL);
    if (rc) return rc;
  }
  return QMC_OK;
}

/* ------------------------------------------------------------------------ */
/* rannyu / setrn / savern (reptation.f:5247-5354): 48-bit LCG in four
 * 12-bit limbs, multiplier limbs (502,1521,4071,2107). */
static const int64_t RN_M[4] = {502, 1521, 4071, 2107};

void orc_setrn(orc_rng* r, const int32_t seed[4]) {
  for (int i = 0; i < 4; i++) r->l[i] = seed[i] % 4096;
  r->l[3] = 2 * (r->l[3] / 2) + 1;
}

void orc_savern(const orc_rng* r, int32_t seed[4]) {
  for (int i = 0; i < 4; i++) seed[i] = (int32_t)r->l[i];
}

double orc_rannyu(orc_rng* r) {
  const double twom12 = 0.000244140625;
  int64_t l1 = r->l[0], l2 = r->l[1], l3 = r->l[2], l4 = r->l[3];
  int64_t m1 = RN_M[0], m2 = RN_M[1], m3 = RN_M[2], m4 = RN_M[3];
  int64_t i1 = l1 * m4 + l2 * m3 + l3 * m2 + l4 * m1;
  int64_t i2 = l2 * m4 + l3 * m3 + l4 * m2;
  int64_t i3 = l3 * m4 + l4 * m3;
  int64_t i4 = l4 * m4;
  l4 = i4 % 4096;
  i3 = i3 + i4 / 4096;
  l3 = i3 % 4096;
  i2 = i2 + i3 / 4096;
  l2 = i2 % 4096;
  l1 = (i1 + i2 / 4096) % 4096;
  r->l[0] = l1; r->l[1] = l2; r->l[2] = l3; r->l[3] = l4;
  return twom12 * (l1 + twom12 * (l2 + twom12 * (l3 + twom12 * l4)));
}

/* the COMMON delta that kinetic's drift cutoff reads (reptation.f:2755) */
int orc_set_delta(orc* o, double delta) {
  o->delta = delta;
  return 0;
}

/* move + metropolis_test replay (reptation.f:4617-4680, 2622-2659), with
 * read_conf's start (reptation.f:2486-2495). */
int orc_vmc_replay(orc* o, const double* x0, int nsteps, double delta,
                   const int32_t seed[4], int bm_order, double* pout, double* ageout,
                   double* ltf_new, double* e_new, double* x_final,
                   int32_t seed_final[4]) {
  const qmc_info* in = &o->info;
  int n = o->nptot, np_ = in->n_props;
  double pi = acos(-1.0);
  double* xo = (double*)malloc(sizeof(double) * 2 * n);
  double* xn = (double*)malloc(sizeof(double) * 2 * n);
  double* go = (double*)malloc(sizeof(double) * 2 * n);
  double* gn = (double*)malloc(sizeof(double) * 2 * n);
  double* po = (double*)calloc(np_, sizeof(double));
  double* pn = (double*)calloc(np_, sizeof(double));
  double ho[2], hn[2], g2ph = 0.0;
  /* the vmc record sets delta, then sigma (reptation.f:740-746, 855-870) */
  o->delta = delta;
  for (int it = 0; it < o->ntypes; it++) {
    o->var[it] = 2.0 * o->hbs2m[it] * delta;
    o->vari[it] = 1.0 / o->var[it];
  }
  memcpy(xn, x0, sizeof(double) * 2 * n);
  int rc = orc_eval(o, xn, pn, gn, hn, &g2ph, NULL);
  if (rc) goto done;
  /* metropolis_test(1): accept */
  memcpy(xo, xn, sizeof(double) * 2 * n);
  memcpy(go, gn, sizeof(double) * 2 * n);
  memcpy(po, pn, sizeof(double) * np_);
  ho[0] = hn[0]; ho[1] = hn[1];
  double age = 0.0;
  po[in->jacc - 1] = 1.0;
  orc_rng rng;
  orc_setrn(&rng, seed);
  for (int s = 0; s < nsteps; s++) {
    double wold = 0.0;
    for (int it = 0; it < o->ntypes; it++) {
      if (o->hbs2m[it] == 0.0) continue;
      double w = 0.0;
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++)
        for (int d = 0; d < 2; d++) {
          double ua, ub;
          if (bm_order == 0) { ua = orc_rannyu(&rng); ub = orc_rannyu(&rng); }
          else { ub = orc_rannyu(&rng); ua = orc_rannyu(&rng); }
          double chi = cos(pi * ua) * sqrt(-2.0 * o->var[it] * log(ub));
          w = w - chi * chi;
          double drift = -o->var[it] * go[2 * ip + d];
          xn[2 * ip + d] = chi + drift + xo[2 * ip + d];
        }
      wold = wold + w * o->vari[it];
    }
    pn[in->jacc - 1] = po[in->jacc - 1];
    rc = orc_eval(o, xn, pn, gn, hn, &g2ph, NULL);
    if (rc) goto done;
    double wnew = 0.0;
    for (int it = 0; it < o->ntypes; it++) {
      if (o->hbs2m[it] == 0.0) continue;
      double w = 0.0;
      for (int ip = o->ipf[it]; ip <= o->ipl[it]; ip++)
        for (int d = 0; d < 2; d++) {
          double drift = -o->var[it] * gn[2 * ip + d];
          double chi = xo[2 * ip + d] - xn[2 * ip + d] - drift;
          w = w - chi * chi;
        }
      wnew = wnew + w * o->vari[it];
    }
    double p = (po[in->jltf - 1] - pn[in->jltf - 1]) * 2 + 0.5 * (wnew - wold);
    p = p > -50.0 ? p : -50.0;
    p = p < 0.0 ? p : 0.0;
    p = exp(p);
    double u = 1.0;
    if (p < 1.0) u = orc_rannyu(&rng);
    ltf_new[s] = pn[in->jltf - 1];
    e_new[s] = pn[in->jetot - 1];
    if (p >= u) {
      age = 0.0;
      memcpy(xo, xn, sizeof(double) * 2 * n);
      memcpy(go, gn, sizeof(double) * 2 * n);
      memcpy(po, pn, sizeof(double) * np_);
      ho[0] = hn[0]; ho[1] = hn[1];
    } else {
      age = age + 1.0;
    }
    po[in->jacc - 1] = p;
    pout[s] = p;
    ageout[s] = age;
  }
  memcpy(x_final, xo, sizeof(double) * 2 * n);
  orc_savern(&rng, seed_final);
done:
  free(xo); free(xn); free(go); free(gn); free(po); free(pn);
  return rc;
}
