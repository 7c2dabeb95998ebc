// qmc_rmc.hip -- reptation QMC moves on the device for a batch of reptiles.
//
// One reptile is the reference's reptation state of one MPI rank: a path of
// ntau beads in a ring of nbuf = ntau + mdelta slots (x_stack, g_stack,
// p_stack, age_stack, w_stack; reptation.f:1766-1802) plus its ends, age,
// path averages and LCG stream.  A move (reptation_move, reptation.f:1897-2113)
// draws a direction and ndelta, grows ndelta beads from the leading end by
// drift-diffusion, each evaluated and passed through a Langevin test, and
// then accepts or rejects the whole slide with the path weight
// (reptation_metropolis_test, reptation.f:1865-1895).  On the device a move is
// k_rmc_begin (directions and ndelta), k_rmc_slots (evaluation slots ordered
// by ndelta), then for idelta = 1..mdelta a proposal kernel, one batched
// evaluation of the reptiles that move bead idelta (the first slots) and a
// bead kernel, then k_rmc_path and k_rmc_finish.  A reptile whose ndelta is
// below idelta is neither evaluated nor draws.
#include "qmc_internal.h"

#include <cstring>

enum { RW_BRA = 0, RW_DTP = 1, RW_ITP = 2, RW_LNG = 3, RW_NDS = 4, RW_N = 5 };
#define RMC_TR 8  // fixed trace fields before the per-bead probabilities

struct RmcK {
  int nr, ntau, nbuf, N, NP, ntypes;
  int jacc, jltf, jetot, jgg, jkin0, jun[2];
  int ipf[2], np[2], mov[2];  // mov: hbs2m != 0 (the types the moves touch)
  double npnorm, delta, etrial, mage, mage_r, var[2], vari[2];
  int ntauskip, mdelta, rejection, ecut, bm_order;
  double value_ecut, half, one, one_nds, mezzo;
};

__device__ __forceinline__ int rslot(int a, int nbuf) { return (a % nbuf + nbuf) % nbuf; }

// fgg (reptation.f:1853-1863)
__device__ __forceinline__ double rmc_fgg(const RmcK& k, double e, double gg) {
  return k.ecut == 0 ? e : k.value_ecut + (e - k.value_ecut) * gg;
}

// -sum_t log(1 - exp(-un_a(t) un_b(t) / (0.5 var_t))) (jnds, reptation.f:2004-2011)
__device__ __forceinline__ double rmc_nds(const RmcK& k, const double* pa, const double* pb) {
  double v = 0.0;
  for (int t = 0; t < k.ntypes; t++)
    if (k.mov[t]) v = v - log(1.0 - exp(-pa[k.jun[t]] * pb[k.jun[t]] / (0.5 * k.var[t])));
  return v;
}

// Reptation direction and number of beads (reptation.f:1921-1927).
// mv[r] = {idir, ndelta, jstart}.
__global__ void k_rmc_begin(RmcK k, const int32_t* __restrict__ ends, int64_t* __restrict__ rng,
                            int32_t* __restrict__ mv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k.nr) return;
  uint64_t st = (uint64_t)rng[r];
  const double u1 = rannyu(st);
  const int idir = (u1 - 0.5 >= 0.0) ? 1 : -1;
  const int nd = 1 + (int)(rannyu(st) * k.mdelta);
  rng[r] = (int64_t)st;
  mv[3 * r] = idir;
  mv[3 * r + 1] = nd;
  mv[3 * r + 2] = idir == 1 ? ends[2 * r + 1] : ends[2 * r];
}

// Evaluation slots of a move: the reptiles ordered by ndelta, largest first,
// so that those moving bead idelta or more hold slots 0..nact[idelta-1]-1
// and each per-bead evaluation covers only them.  One workgroup; the order
// inside a group of equal ndelta is whatever the LDS atomics give (a
// reptile's evaluation does not depend on its slot).
#define RMC_SLOT_THREADS 1024
__global__ __launch_bounds__(RMC_SLOT_THREADS) void k_rmc_slots(RmcK k, const int32_t* __restrict__ mv,
                                                              int32_t* __restrict__ slot,
                                                              int32_t* __restrict__ nact) {
  __shared__ int cnt[RMC_MAX_MDELTA + 2], base[RMC_MAX_MDELTA + 2], fill[RMC_MAX_MDELTA + 2];
  const int tid = threadIdx.x;
  for (int b = tid; b < RMC_MAX_MDELTA + 2; b += RMC_SLOT_THREADS) cnt[b] = base[b] = fill[b] = 0;
  __syncthreads();
  for (int r = tid; r < k.nr; r += RMC_SLOT_THREADS) atomicAdd(&cnt[mv[3 * r + 1]], 1);
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int b = k.mdelta; b >= 1; b--) {
      base[b] = acc;
      acc += cnt[b];
      nact[b - 1] = acc;  // reptiles with ndelta >= b
    }
  }
  __syncthreads();
  for (int r = tid; r < k.nr; r += RMC_SLOT_THREADS) {
    const int b = mv[3 * r + 1];
    slot[r] = base[b] + atomicAdd(&fill[b], 1);
  }
}

// Bead proposal idelta (reptation.f:1929-1955): x_new = chi - var g_j + x_j
// from bead j, with wold = -sum chi^2/var, dold = sum drift^2/var and
// lng = -sum (chi + drift)^2/var.  sc[r] = {wold, dold, lng}.
// One wavefront per reptile.  The reference draws two uniforms per moved
// coordinate in coordinate order, so the coordinate with ordinal o among the
// moved ones takes draws 2o+1 and 2o+2 of the reptile's stream: each lane
// jumps there with a^(2o+1).  The per-coordinate terms go to LDS and three
// lanes add them in the reference's order.
__global__ __launch_bounds__(QMC_BLOCK) void k_rmc_propose(
    RmcK k, int idelta, const double* __restrict__ X, const double* __restrict__ G,
    const int32_t* __restrict__ mv, const int32_t* __restrict__ slot, int64_t* __restrict__ rng,
    double* __restrict__ xn, double* __restrict__ sc, const double* __restrict__ g2,
    double* __restrict__ g2n) {
  extern __shared__ double rp_sm[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = blockIdx.x * (QMC_BLOCK / 64) + wv;
  // every wavefront reaches the barrier below: no early return
  const bool live = r < k.nr;
  const int idir = live ? mv[3 * r] : 0, nd = live ? mv[3 * r + 1] : 0, js = live ? mv[3 * r + 2] : 0;
  const bool prop = live && idelta <= nd;
  const int N = k.N, n2 = 2 * N;
  // the proposal goes to the reptile's evaluation slot: the reptiles still
  // moving a bead hold the first slots (k_rmc_slots)
  const int sl = prop ? slot[r] : 0;
  double* xo = xn + (int64_t)sl * n2;
  if (prop && lane == 0) g2n[sl] = g2[r];
  const int j = rslot(js + idir * (idelta - 1), k.nbuf);
  const double* xj = X + ((int64_t)r * k.nbuf + j) * n2;
  const double* gj = G + ((int64_t)r * k.nbuf + j) * n2;
  const uint64_t st0 = prop ? (uint64_t)rng[r] : 0;
  const double pi = 3.141592653589793;
  const uint64_t msk = (1ULL << 48) - 1ULL, a = 34522712143931ULL;
  const bool two = k.ntypes > 1;
  const int m0 = k.mov[0] ? 2 * k.np[0] : 0;             // moved coordinates of type 0
  const int nmov = m0 + (two && k.mov[1] ? 2 * k.np[1] : 0);
  double* tw = rp_sm + wv * 3 * n2;  // chi^2, drift^2, (chi + drift)^2 per coordinate
  double* td = tw + n2;
  double* tg = td + n2;
  for (int e = lane; prop && e < n2; e += 64) {
    const int t = two && e >= 2 * k.ipf[1] ? 1 : 0;
    if (!(t ? k.mov[1] : k.mov[0])) continue;
    const int o = t ? m0 + e - 2 * k.ipf[1] : e - 2 * k.ipf[0];
    const double var = t ? k.var[1] : k.var[0];
    const uint64_t x1 = (lcg_pow(2 * (uint64_t)o + 1) * st0) & msk, x2 = (a * x1) & msk;
    const double u1 = (double)x1 * 0x1p-48, u2 = (double)x2 * 0x1p-48;
    const double ua = k.bm_order == 0 ? u1 : u2, ub = k.bm_order == 0 ? u2 : u1;
    const double drift = -var * gj[e];
    const double chi = cos(pi * ua) * sqrt(-2.0 * var * log(ub));
    xo[e] = chi + drift + xj[e];
    const double cd = chi + drift;
    tw[e] = chi * chi;
    td[e] = drift * drift;
    tg[e] = cd * cd;
  }
  __syncthreads();
  if (prop && lane < 3) {
    const double* tq = lane == 0 ? tw : (lane == 1 ? td : tg);
    double acc = 0.0;
    for (int t = 0; t < k.ntypes; t++) {
      if (!(t ? k.mov[1] : k.mov[0])) continue;
      const int e0 = 2 * (t ? k.ipf[1] : k.ipf[0]), e1 = e0 + 2 * (t ? k.np[1] : k.np[0]);
      double v = 0.0;
      if (lane == 1)
        for (int e = e0; e < e1; e++) v = v + tq[e];
      else
        for (int e = e0; e < e1; e++) v = v - tq[e];
      acc = acc + v * (t ? k.vari[1] : k.vari[0]);
    }
    sc[3 * sl + (lane == 0 ? 0 : (lane == 1 ? 1 : 2))] = acc;
    if (lane == 0) rng[r] = (int64_t)((lcg_pow(2 * (uint64_t)nmov) * st0) & msk);
  }
}

// Langevin test of bead idelta and the ring update (reptation.f:1956-2104).
// One wavefront per reptile.
__global__ __launch_bounds__(QMC_BLOCK) void k_rmc_step(
    RmcK k, int idelta, double* __restrict__ X, double* __restrict__ G, double* __restrict__ P,
    double* __restrict__ AGE, double* __restrict__ W, const int32_t* __restrict__ mv,
    const int32_t* __restrict__ slot, int64_t* __restrict__ rng, const double* __restrict__ xn,
    const double* __restrict__ gn,
    const double* __restrict__ pn, const double* __restrict__ sc, double* __restrict__ g2,
    const double* __restrict__ g2n, const int32_t* __restrict__ stn, int32_t* __restrict__ sticky,
    double* __restrict__ trace) {
  const int r = blockIdx.x * (QMC_BLOCK / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= k.nr) return;
  const int idir = mv[3 * r], nd = mv[3 * r + 1], js = mv[3 * r + 2];
  const int rowlen = RMC_TR + k.mdelta;
  if (idelta > nd) {
    if (lane == 0 && trace) trace[(int64_t)r * rowlen + RMC_TR + idelta - 1] = -1.0;
    return;
  }
  const int N = k.N, NP = k.NP, nbuf = k.nbuf;
  const int j = rslot(js + idir * (idelta - 1), nbuf);
  const int jn = rslot(js + idir * idelta, nbuf);
  double* xj = X + ((int64_t)r * nbuf + j) * 2 * N;
  double* gj = G + ((int64_t)r * nbuf + j) * 2 * N;
  double* pj = P + ((int64_t)r * nbuf + j) * NP;
  double* xq = X + ((int64_t)r * nbuf + jn) * 2 * N;
  double* gq = G + ((int64_t)r * nbuf + jn) * 2 * N;
  double* pq = P + ((int64_t)r * nbuf + jn) * NP;
  const int sl = slot[r];  // the reptile's evaluation slot
  const double* xr = xn + (int64_t)sl * 2 * N;
  const double* gr = gn + (int64_t)sl * 2 * N;
  const double* pr = pn + (int64_t)sl * NP;
  // reverse-move weight (reptation.f:1958-1970)
  double wt[2] = {0.0, 0.0};
  for (int ip = lane; ip < N; ip += 64) {
    const int t = ip >= k.ipf[1] && k.ntypes > 1 ? 1 : 0;
    for (int dm = 0; dm < 2; dm++) {
      const double drift = -k.var[t] * gr[2 * ip + dm];
      const double chi = xj[2 * ip + dm] - xr[2 * ip + dm] - drift;
      if (t == 0) wt[0] = wt[0] - chi * chi;
      else wt[1] = wt[1] - chi * chi;
    }
  }
  wt[0] = wave_sum(wt[0]);
  wt[1] = wave_sum(wt[1]);
  int acc = 0;
  double p = 0.0;
  if (lane == 0) {
    double* Wr = W + (int64_t)r * nbuf * RW_N;
    double* wj = Wr + j * RW_N;
    double* wq = Wr + jn * RW_N;
    const double wold = sc[3 * sl], dold = sc[3 * sl + 1];
    double lng = sc[3 * sl + 2];
    double wnew = 0.0;
    for (int t = 0; t < k.ntypes; t++)
      if (k.mov[t]) wnew = wnew + wt[t] * k.vari[t];
    p = 2.0 * (pj[k.jltf] - pr[k.jltf]) + 0.5 * (wnew - wold);
    p = fmin(0.0, fmax(-50.0, p));
    p = exp(p);
    if (k.rejection == 0) p = 1.0;
    else if (AGE[(int64_t)r * nbuf + j] == k.mage) p = 1.0;
    const double wg = k.delta * k.npnorm * (pj[k.jkin0] + pr[k.jkin0]);
    uint64_t st = (uint64_t)rng[r];
    const double u = rannyu(st);
    rng[r] = (int64_t)st;
    const double fj = rmc_fgg(k, pj[k.jetot], pj[k.jgg]);
    if (p < u) {
      // rejected: bead j is repeated (reptation.f:1992-2045)
      const double bra = -(k.delta * k.npnorm * (fj - k.etrial));
      const double nds = rmc_nds(k, pj, pj);
      if (idir == 1) {
        wj[RW_DTP] = 0.5 * dold; wj[RW_BRA] = bra; wq[RW_ITP] = 0.5 * dold;
        wj[RW_LNG] = 0.5 * wg; wj[RW_NDS] = nds;
      } else {
        wq[RW_DTP] = 0.5 * dold; wq[RW_BRA] = bra; wj[RW_ITP] = 0.5 * dold;
        wq[RW_LNG] = 0.5 * wg; wq[RW_NDS] = nds;
      }
      AGE[(int64_t)r * nbuf + jn] = AGE[(int64_t)r * nbuf + j] + 1.0;
    } else {
      // accepted (reptation.f:2046-2101)
      acc = 1;
      lng = lng + wg;
      const double bra = -(0.5 * k.delta * k.npnorm *
                           ((fj + rmc_fgg(k, pr[k.jetot], pr[k.jgg])) - 2.0 * k.etrial));
      const double nds = rmc_nds(k, pj, pr);
      if (idir == 1) {
        wj[RW_DTP] = 0.5 * wold; wj[RW_BRA] = bra; wq[RW_ITP] = 0.5 * wnew;
        wj[RW_LNG] = 0.5 * lng; wj[RW_NDS] = nds;
      } else {
        wq[RW_DTP] = 0.5 * wnew; wq[RW_BRA] = bra; wj[RW_ITP] = 0.5 * wold;
        wq[RW_LNG] = 0.5 * lng; wq[RW_NDS] = nds;
      }
      AGE[(int64_t)r * nbuf + jn] = 0.0;
    }
    g2[r] = g2n[sl];
    sticky[r] |= stn[sl];
    if (trace) trace[(int64_t)r * rowlen + RMC_TR + idelta - 1] = p;
  }
  acc = __shfl(acc, 0, 64);
  p = __shfl(p, 0, 64);
  const double* xs = acc ? xr : xj;
  const double* gs = acc ? gr : gj;
  const double* ps = acc ? pr : pj;
  for (int e = lane; e < 2 * N; e += 64) {
    xq[e] = xs[e];
    gq[e] = gs[e];
  }
  for (int e = lane; e < NP; e += 64) pq[e] = (e == k.jacc) ? p : ps[e];
}

// path_props scalars (reptation.f:1566-1577) for the window starting at slot
// kf: mean over beads ntauskip+1..ntau-ntauskip, E from the two end beads.
// One wavefront per (reptile, 64 properties); each lane adds its property's
// beads in the reference's order, with RMC_PB loads in flight at a time.
// With mv the window is the one the move slides to (kf, kl from ends and the
// move), else slots 0..ntau-1.
#define RMC_PB 16
__global__ __launch_bounds__(64) void k_rmc_path(RmcK k, const double* __restrict__ P,
                                                 const int32_t* __restrict__ ends,
                                                 const int32_t* __restrict__ mv,
                                                 double* __restrict__ out) {
  const int r = blockIdx.x, ip = blockIdx.y * 64 + threadIdx.x;
  const int nbuf = k.nbuf, NP = k.NP;
  if (ip >= NP) return;
  int kf = 0, kl = k.ntau - 1;
  if (mv) {
    const int idir = mv[3 * r], nd = mv[3 * r + 1];
    kf = rslot(ends[2 * r] + idir * nd, nbuf);
    kl = rslot(ends[2 * r + 1] + idir * nd, nbuf);
  }
  const double* Pr = P + (int64_t)r * nbuf * NP;
  const int nb = k.ntau - 2 * k.ntauskip;
  int sl = rslot(kf + k.ntauskip, nbuf);
  double v = 0.0;
  for (int i0 = 0; i0 < nb; i0 += RMC_PB) {
    double t[RMC_PB];
#pragma unroll
    for (int q = 0; q < RMC_PB; q++) {
      t[q] = i0 + q < nb ? Pr[(int64_t)sl * NP + ip] : 0.0;
      sl = sl + 1 == nbuf ? 0 : sl + 1;
    }
#pragma unroll
    for (int q = 0; q < RMC_PB; q++)
      if (i0 + q < nb) v = v + t[q];
  }
  v = v / (double)nb;
  if (ip == k.jetot) v = 0.5 * (Pr[(int64_t)kl * NP + k.jetot] + Pr[(int64_t)kf * NP + k.jetot]);
  out[(int64_t)r * NP + ip] = v;
}

// End of the move: path weight (reptation.f:2105-2160),
// reptation_metropolis_test and averages(2), with k_rmc_path's averages of
// the slid window.
// One wavefront per reptile.
__global__ __launch_bounds__(64) void k_rmc_finish(
    RmcK k, const double* __restrict__ P, const double* __restrict__ W, int32_t* __restrict__ ends,
    const int32_t* __restrict__ mv, int64_t* __restrict__ rng, double* __restrict__ rage,
    double* __restrict__ pp_old, const double* __restrict__ ppn, double* __restrict__ blk,
    double* __restrict__ trace) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const int nbuf = k.nbuf, NP = k.NP;
  const int idir = mv[3 * r], nd = mv[3 * r + 1];
  const int jf = ends[2 * r], jl = ends[2 * r + 1];
  const int kf = rslot(jf + idir * nd, nbuf), kl = rslot(jl + idir * nd, nbuf);
  const double* Pr = P + (int64_t)r * nbuf * NP;
  const double* Wr = W + (int64_t)r * nbuf * RW_N;
  const double* pn = ppn + (int64_t)r * NP;  // k_rmc_path's averages of the slid window
  __shared__ int s_acc;
  if (lane == 0) {
    const double m = k.mezzo, h = k.half, o = k.one, on = k.one_nds;
    double wn = (-Pr[(int64_t)kf * NP + k.jltf] * (1 + 2 * m) - Pr[(int64_t)kl * NP + k.jltf] * (1 - 2 * m) +
                 Pr[(int64_t)jf * NP + k.jltf] * (1 + 2 * m) + Pr[(int64_t)jl * NP + k.jltf] * (1 - 2 * m));
    for (int id = 1; id <= nd; id++) {
      if (idir == 1) {
        const double* a = Wr + rslot(jf + id - 1, nbuf) * RW_N;
        const double* b = Wr + rslot(jf + id, nbuf) * RW_N;
        wn = wn - a[RW_BRA] - a[RW_DTP] * (h + m) + b[RW_ITP] * (o + h + m) - a[RW_LNG] * o +
             a[RW_NDS] * on;
      } else {
        const double* a = Wr + rslot(jl - (id - 1), nbuf) * RW_N;
        const double* b = Wr + rslot(jl - id, nbuf) * RW_N;
        wn = wn - b[RW_BRA] + b[RW_DTP] * (o + h - m) - a[RW_ITP] * (h - m) - b[RW_LNG] * o +
             b[RW_NDS] * on;
      }
    }
    for (int id = 1; id <= nd; id++) {
      if (idir == 1) {
        const double* a = Wr + rslot(jl + id - 1, nbuf) * RW_N;
        const double* b = Wr + rslot(jl + id, nbuf) * RW_N;
        wn = wn + a[RW_BRA] - a[RW_DTP] * (o + h - m) + b[RW_ITP] * (h - m) + a[RW_LNG] * o -
             a[RW_NDS] * on;
      } else {
        const double* a = Wr + rslot(jf - (id - 1), nbuf) * RW_N;
        const double* b = Wr + rslot(jf - id, nbuf) * RW_N;
        wn = wn + b[RW_BRA] + b[RW_DTP] * (h + m) - a[RW_ITP] * (o + h + m) + b[RW_LNG] * o -
             b[RW_NDS] * on;
      }
    }
    // reptation_metropolis_test (reptation.f:1865-1895)
    double p;
    if (rage[r] == k.mage_r) p = 1.0;
    else p = exp(fmin(0.0, fmax(-50.0, wn)));
    uint64_t st = (uint64_t)rng[r];
    const double u = rannyu(st);
    rng[r] = (int64_t)st;
    const int acc = p >= u;
    if (acc) {
      ends[2 * r] = kf;
      ends[2 * r + 1] = kl;
      rage[r] = 0.0;
    } else {
      rage[r] = rage[r] + 1.0;
    }
    s_acc = acc;
    if (trace) {
      double* tr = trace + (int64_t)r * (RMC_TR + k.mdelta);
      tr[0] = idir;
      tr[1] = nd;
      tr[2] = wn;
      tr[3] = rage[r];
    }
  }
  __syncthreads();
  const int acc = s_acc;
  double* po = pp_old + (int64_t)r * NP;
  double* b = blk ? blk + (int64_t)r * NP : nullptr;
  for (int ip = lane; ip < NP; ip += 64) {
    const double v = acc ? pn[ip] : po[ip];
    po[ip] = v;
    if (b) b[ip] += v;
  }
  if (trace && lane == 0) {
    double* tr = trace + (int64_t)r * (RMC_TR + k.mdelta);
    tr[4] = acc ? pn[k.jetot] : po[k.jetot];
    tr[5] = acc ? pn[k.jltf] : po[k.jltf];
    tr[6] = acc ? pn[k.jkin0] : po[k.jkin0];
    tr[7] = 0.0;
  }
}

// read_path: bead b of every reptile to the staging row, and the evaluated
// bead back into the ring, accepted as read_conf does (age 0, p(jacc) = 1).
__global__ void k_rmc_gather(RmcK k, int b, const double* __restrict__ X, double* __restrict__ xn) {
  const int r = blockIdx.x;
  for (int e = threadIdx.x; e < 2 * k.N; e += blockDim.x)
    xn[(int64_t)r * 2 * k.N + e] = X[((int64_t)r * k.nbuf + b) * 2 * k.N + e];
}

__global__ void k_rmc_scatter(RmcK k, int b, double* __restrict__ G, double* __restrict__ P,
                              double* __restrict__ AGE, const double* __restrict__ gn,
                              const double* __restrict__ pn, const int32_t* __restrict__ stn,
                              int32_t* __restrict__ sticky) {
  const int r = blockIdx.x;
  for (int e = threadIdx.x; e < 2 * k.N; e += blockDim.x)
    G[((int64_t)r * k.nbuf + b) * 2 * k.N + e] = gn[(int64_t)r * 2 * k.N + e];
  for (int e = threadIdx.x; e < k.NP; e += blockDim.x)
    P[((int64_t)r * k.nbuf + b) * k.NP + e] = e == k.jacc ? 1.0 : pn[(int64_t)r * k.NP + e];
  if (threadIdx.x == 0) {
    AGE[(int64_t)r * k.nbuf + b] = 0.0;
    sticky[r] |= stn[r];
  }
}

// read_path after the beads and k_rmc_path's averages of slots 0..ntau-1:
// the link weights and Langevin probabilities of every link (get_w_stack,
// reptation.f:1804-1851, and reptation.f:1789-1798).  One wavefront per
// reptile.
__global__ __launch_bounds__(64) void k_rmc_links(RmcK k, const double* __restrict__ X,
                                                  const double* __restrict__ G, double* __restrict__ P,
                                                  double* __restrict__ W, int32_t* __restrict__ ends,
                                                  double* __restrict__ rage) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = k.N, NP = k.NP, nbuf = k.nbuf;
  double* Pr = P + (int64_t)r * nbuf * NP;
  double* Wr = W + (int64_t)r * nbuf * RW_N;
  for (int i = 1; i < k.ntau; i++) {
    const int j = i - 1, jn = i;
    const double* xj = X + ((int64_t)r * nbuf + j) * 2 * N;
    const double* xq = X + ((int64_t)r * nbuf + jn) * 2 * N;
    const double* gj = G + ((int64_t)r * nbuf + j) * 2 * N;
    const double* gq = G + ((int64_t)r * nbuf + jn) * 2 * N;
    double wf[2] = {0, 0}, wb[2] = {0, 0}, wg[2] = {0, 0};
    for (int ip = lane; ip < N; ip += 64) {
      const int t = ip >= k.ipf[1] && k.ntypes > 1 ? 1 : 0;
      for (int dm = 0; dm < 2; dm++) {
        const double a = xj[2 * ip + dm], c = xq[2 * ip + dm];
        const double df = -k.var[t] * gj[2 * ip + dm];
        const double db = -k.var[t] * gq[2 * ip + dm];
        const double eg = (a - c) * (a - c);
        const double ef = (a + df - c) * (a + df - c);
        const double eb = (a - db - c) * (a - db - c);
        wg[t] = wg[t] - eg;
        wf[t] = wf[t] - ef;
        wb[t] = wb[t] - eb;
      }
    }
    for (int t = 0; t < 2; t++) {
      wg[t] = wave_sum(wg[t]);
      wf[t] = wave_sum(wf[t]);
      wb[t] = wave_sum(wb[t]);
    }
    if (lane == 0) {
      const double* pj = Pr + (int64_t)j * NP;
      double* pq = Pr + (int64_t)jn * NP;
      double wnew = 0.0, wold = 0.0, lng = 0.0;
      for (int t = 0; t < k.ntypes; t++) {
        if (!k.mov[t]) continue;
        wnew = wnew + wb[t] * k.vari[t];
        wold = wold + wf[t] * k.vari[t];
        lng = lng + wg[t] * k.vari[t];
      }
      lng = lng + k.delta * k.npnorm * (pj[k.jkin0] + pq[k.jkin0]);
      double* wj = Wr + j * RW_N;
      double* wq = Wr + jn * RW_N;
      wj[RW_LNG] = 0.5 * lng;
      wj[RW_DTP] = 0.5 * wold;
      wj[RW_BRA] = -(0.5 * k.delta * k.npnorm *
                     (rmc_fgg(k, pj[k.jetot], pj[k.jgg]) + rmc_fgg(k, pq[k.jetot], pq[k.jgg]) -
                      2.0 * k.etrial));
      wq[RW_ITP] = 0.5 * wnew;
      wj[RW_NDS] = rmc_nds(k, pj, pq);
      double p = 2.0 * (pj[k.jltf] - pq[k.jltf]) + 0.5 * (wnew - wold);
      p = fmin(0.0, fmax(-50.0, p));
      pq[k.jacc] = exp(p);
    }
    __syncthreads();
  }
  if (lane == 0) {
    ends[2 * r] = 0;
    ends[2 * r + 1] = k.ntau - 1;
    rage[r] = 0.0;
  }
}

static int ensure_rmc(qmc_ctx* c, int nr) {
  if (nr <= c->rmc_nr) return QMC_OK;
  (void)hipFree(c->d_rmv);
  (void)hipFree(c->d_rsc);
  (void)hipFree(c->d_rslot);
  c->d_rmv = nullptr;
  c->d_rsc = nullptr;
  c->d_rslot = nullptr;
  const size_t NP = c->info.n_props;
  HIPCHK(hipMalloc(&c->d_rmv, sizeof(int32_t) * 3 * nr));
  // sc[3], g2n[1], path averages [NP]
  HIPCHK(hipMalloc(&c->d_rsc, sizeof(double) * (4 + NP) * nr));
  // evaluation slot per reptile, then the counts per bead of a move
  HIPCHK(hipMalloc(&c->d_rslot, sizeof(int32_t) * (nr + RMC_MAX_MDELTA)));
  c->rmc_nr = nr;
  return QMC_OK;
}

void free_rmc(qmc_ctx* c) {
  (void)hipFree(c->d_rmv);
  (void)hipFree(c->d_rsc);
  (void)hipFree(c->d_rslot);
  c->d_rmv = nullptr;
  c->d_rsc = nullptr;
  c->d_rslot = nullptr;
  c->rmc_nr = 0;
}

static int rmc_setup(qmc_ctx* c, const qmc_reptiles* r, const qmc_rmc_params* prm, RmcK& k) {
  if (!c || !r || !prm || r->nr < 0) return arg_fail("bad reptile set");
  const DevSys& s = c->s;
  if (r->nr > 0 && (!r->x || !r->g || !r->props || !r->age || !r->w || !r->ends || !r->rage ||
                    !r->pp_old || !r->grad2_ph || !r->rng))
    return arg_fail("null reptile buffer");
  if (!(prm->delta > 0.0)) return arg_fail("delta must be positive");
  if (prm->mdelta < 1 || r->ntau < 2 || r->nbuf != r->ntau + prm->mdelta)
    return arg_fail("reptile needs ntau >= 2, mdelta >= 1 and nbuf = ntau + mdelta");
  if (prm->mdelta > RMC_MAX_MDELTA) return arg_fail("mdelta above 64", QMC_ERR_LIMIT);
  if (prm->ntauskip < 0 || r->ntau - 2 * prm->ntauskip < 1)
    return arg_fail("ntauskip leaves no bead for the path averages");
  for (int t = 0; t < s.ntypes; t++)
    if (s.np[t] > 0 && s.hbs2m[t] == 0.0)
      return arg_fail("reptation needs every populated type to move (hbs2m != 0)");
  // the rmc record sets the COMMON delta before sigma (reptation.f:795-805)
  c->delta = prm->delta;
  // every byte set (padding included): the captured move is keyed on them
  memset(&k, 0, sizeof k);
  k.nr = r->nr;
  k.ntau = r->ntau;
  k.nbuf = r->nbuf;
  k.N = s.nptot;
  k.NP = c->info.n_props;
  k.ntypes = s.ntypes;
  k.jacc = c->info.jacc - 1;
  k.jltf = c->info.jltf - 1;
  k.jetot = c->info.jetot - 1;
  k.jgg = c->info.jgg - 1;
  k.jkin0 = c->info.jkin0 - 1;
  for (int t = 0; t < 2; t++) {
    k.jun[t] = c->info.jun[t] - 1;
    k.ipf[t] = s.ipf[t];
    k.np[t] = t < s.ntypes ? s.np[t] : 0;
    // sigma (reptation.f:855-870)
    k.mov[t] = t < s.ntypes && s.hbs2m[t] != 0.0;
    const double var = k.mov[t] ? 2.0 * s.hbs2m[t] * prm->delta : 1.0;
    k.var[t] = var;
    k.vari[t] = 1.0 / var;
  }
  if (s.ntypes < 2) k.ipf[1] = s.nptot;
  k.npnorm = c->info.npnorm;
  k.delta = prm->delta;
  k.etrial = prm->etrial;
  k.mage = prm->mage;
  k.mage_r = prm->mage_r;
  k.ntauskip = prm->ntauskip;
  k.mdelta = prm->mdelta;
  k.rejection = prm->rejection;
  k.ecut = prm->ecut;
  k.value_ecut = prm->value_ecut;
  k.bm_order = prm->bm_order;
  // propagator coefficients (reptation.f:1904-1919)
  k.half = 0.5;
  k.one = 0.0;
  k.one_nds = 0.0;
  k.mezzo = 0.0;
  if (prm->wpiu != 0) {
    k.mezzo = 0.5;
  } else if (prm->gstorto != 0) {
    k.half = 0.0;
    k.one = 1.0;
    if (prm->nodalaction != 0) k.one_nds = 1.0;
  }
  return QMC_OK;
}

extern "C" int qmc_rmc_start(qmc_ctx* c, const qmc_reptiles* r, const qmc_rmc_params* prm,
                             void* stream) {
  RmcK k;
  int rc = rmc_setup(c, r, prm, k);
  if (rc) return rc;
  if (r->nr == 0) return QMC_OK;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  rc = ensure_vmc_scratch(c, r->nr);
  if (rc) return rc;
  rc = ensure_rmc(c, r->nr);
  if (rc) return rc;
  const int nr = r->nr;
  HIPCHK(hipMemsetAsync(c->d_sticky, 0, sizeof(int32_t) * nr, st));
  for (int b = 0; b < r->ntau; b++) {
    hipLaunchKernelGGL(k_rmc_gather, dim3(nr), dim3(64), 0, st, k, b, r->x, c->d_xn);
    rc = launch_eval(c, nr, c->d_xn, c->d_pn, c->d_gn, c->d_hn, r->grad2_ph, c->d_stn, st,
                     prm->delta);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rmc_scatter, dim3(nr), dim3(64), 0, st, k, b, r->g, r->props, r->age,
                       c->d_gn, c->d_pn, c->d_stn, c->d_sticky);
  }
  hipLaunchKernelGGL(k_rmc_path, dim3(nr, (k.NP + 63) / 64), dim3(64), 0, st, k, r->props,
                     r->ends, nullptr, r->pp_old);
  hipLaunchKernelGGL(k_rmc_links, dim3(nr), dim3(64), 0, st, k, r->x, r->g, r->props, r->w,
                     r->ends, r->rage);
  HIPCHK(hipGetLastError());
  return QMC_OK;
}

extern "C" int qmc_rmc_sweep(qmc_ctx* c, const qmc_reptiles* r, const qmc_rmc_params* prm,
                             int nsteps, double* trace, void* stream) {
  RmcK k;
  int rc = rmc_setup(c, r, prm, k);
  if (rc) return rc;
  if (r->nr == 0 || nsteps <= 0) return QMC_OK;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  rc = ensure_vmc_scratch(c, r->nr);
  if (rc) return rc;
  rc = ensure_rmc(c, r->nr);
  if (rc) return rc;
  rc = ensure_ws(c, r->nr);
  if (rc) return rc;
  const int nr = r->nr;
  double* sc = c->d_rsc;
  double* g2n = c->d_rsc + 3 * nr;
  double* ppn = c->d_rsc + 4 * nr;
  const int wpb = QMC_BLOCK / 64;
  const int64_t rowlen = RMC_TR + prm->mdelta;
  // one move of every reptile on stream q (trace row tr, or none)
  auto move = [&](hipStream_t q, double* tr) -> int {
    hipLaunchKernelGGL(k_rmc_begin, dim3((nr + 63) / 64), dim3(64), 0, q, k, r->ends, r->rng, c->d_rmv);
    hipLaunchKernelGGL(k_rmc_slots, dim3(1), dim3(RMC_SLOT_THREADS), 0, q, k, c->d_rmv, c->d_rslot,
                       c->d_rslot + nr);
    for (int id = 1; id <= prm->mdelta; id++) {
      hipLaunchKernelGGL(k_rmc_propose, dim3((nr + wpb - 1) / wpb), dim3(QMC_BLOCK),
                         sizeof(double) * 3 * 2 * k.N * wpb, q, k, id, r->x, r->g, c->d_rmv,
                         c->d_rslot, r->rng, c->d_xn, sc, r->grad2_ph, g2n);
      // only the reptiles moving bead id are evaluated (slots 0..nact-1)
      const int rc2 = launch_eval(c, nr, c->d_xn, c->d_pn, c->d_gn, c->d_hn, g2n, c->d_stn, q, prm->delta,
                                  QMC_PREC_FP64, c->d_rslot + nr + id - 1);
      if (rc2) return rc2;
      hipLaunchKernelGGL(k_rmc_step, dim3((nr + wpb - 1) / wpb), dim3(QMC_BLOCK), 0, q, k, id, r->x, r->g,
                         r->props, r->age, r->w, c->d_rmv, c->d_rslot, r->rng, c->d_xn, c->d_gn, c->d_pn, sc,
                         r->grad2_ph, g2n, c->d_stn, c->d_sticky, tr);
    }
    hipLaunchKernelGGL(k_rmc_path, dim3(nr, (k.NP + 63) / 64), dim3(64), 0, q, k, r->props, r->ends,
                       c->d_rmv, ppn);
    hipLaunchKernelGGL(k_rmc_finish, dim3(nr), dim3(64), 0, q, k, r->props, r->w, r->ends, c->d_rmv,
                       r->rng, r->rage, r->pp_old, ppn, r->blk, tr);
    return QMC_OK;
  };
  const bool graph = c->use_graph && !trace && !c->timing && nsteps > 1;
  if (!graph) {
    for (int s = 0; s < nsteps; s++) {
      rc = move(st, trace ? trace + (int64_t)s * nr * rowlen : nullptr);
      if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    return QMC_OK;
  }
  // The move as a captured graph, replayed nsteps times, as the VMC step:
  // rebuilt whenever anything its kernels were given changes (buffers, the
  // move's parameters, the g-vectors).  Captured on the context's own
  // stream, launched on st.
  std::vector<uint64_t> key = {(uint64_t)nr,
                               (uint64_t)(uintptr_t)r->x,       (uint64_t)(uintptr_t)r->g,
                               (uint64_t)(uintptr_t)r->props,   (uint64_t)(uintptr_t)r->age,
                               (uint64_t)(uintptr_t)r->w,       (uint64_t)(uintptr_t)r->ends,
                               (uint64_t)(uintptr_t)r->rage,    (uint64_t)(uintptr_t)r->pp_old,
                               (uint64_t)(uintptr_t)r->grad2_ph, (uint64_t)(uintptr_t)r->rng,
                               (uint64_t)(uintptr_t)r->blk,     (uint64_t)(uintptr_t)c->d_xn,
                               (uint64_t)(uintptr_t)c->d_rmv,   (uint64_t)(uintptr_t)c->d_rslot,
                               (uint64_t)(uintptr_t)c->d_rsc,   (uint64_t)(uintptr_t)c->d_ws,
                               (uint64_t)(uintptr_t)c->d_F,     (uint64_t)(uintptr_t)c->d_A,
                               (uint64_t)(uintptr_t)c->s.gvec,  (uint64_t)c->s.ng};
  {
    uint64_t words[(sizeof(RmcK) + 7) / 8] = {};
    memcpy(words, &k, sizeof k);
    for (uint64_t wd : words) key.push_back(wd);
  }
  if (!c->rmc_exec || key != c->rmc_key) {
    if (c->rmc_exec) (void)hipGraphExecDestroy(c->rmc_exec);
    c->rmc_exec = nullptr;
    c->rmc_key.clear();
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = move(c->stream, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    const hipError_t e2 = hipGraphInstantiate(&c->rmc_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      c->rmc_exec = nullptr;
      return hip_fail(e2, "hipGraphInstantiate");
    }
    c->rmc_key = key;
  }
  for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(c->rmc_exec, st));
  HIPCHK(hipGetLastError());
  return QMC_OK;
}
