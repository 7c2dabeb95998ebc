// qmc_gj.hip -- determinant path for 33..64 particles per spin.
//
// k_gj: F_b = A^-1 (G_b A) (zv of reptation.f:10182-10194, G_b = diag(i g^b))
// and log|det A| by one Gauss-Jordan solve of A X = [G_x A | G_y A] held in
// registers.  Lane = row (orbital).  The 3m columns [A | B_x | B_y] are dealt
// to the 8 wavefronts cyclically: column c sits in slot c / 8 of wavefront
// c % 8, so step k's pivot column is a compile-time slot of one wavefront
// and the A columns already eliminated are whole slots that later steps skip.
//
// Each step picks the pivot row of column k among the rows not yet used by
// the largest |re|+|im| (zgefa's criterion, reptation.f:9286-9298; rows whose
// magnitudes agree to 13 bits may be taken in another order, which changes
// rounding only).  Rows are never swapped and never scaled during the sweep:
// every other row takes v -= (f / piv) u with u the raw pivot row, and the
// pivot row is divided by its pivot once at the end, so row p of step k ends
// up holding row k of X.  Each wavefront takes the pivot row's values for
// its own columns from the pivot lane with v_readlane.  The wavefront that
// owns column k+1 finishes its update and publishes the next pivot (column
// values, row, 1/piv) before the step's single barrier, so the search never
// sits between two barriers.
//
// k_fgrad: drift and the diagonal Laplacian terms (reptation.f:10195-10240)
// from F: d_j = F_b(j,j) and diag(A^-1 d_a d_b A)_j = sum_k F_a(j,k) F_b(k,j)
// (A^-1 G_a G_b A = F_a F_b), with the backflow Jacobian regenerated from
// the eta table.
#include "qmc_dev.h"

#define GJ_WAVES 8
#define GJ_CW 24
#define GJ_MP 64
#define GJ_THREADS (GJ_WAVES * 64)

// Empty statement that takes a value in registers: the arithmetic producing
// it happens before this point and is not sunk past it.
__device__ __forceinline__ void gj_pin(zd& a) { __asm__ volatile("" : "+v"(a.x), "+v"(a.y)); }

struct __attribute__((aligned(16))) GjShared {
  // pivot column of the step, double-buffered
  __attribute__((aligned(16))) zd mbuf[2][GJ_MP];
  zd pinv[GJ_MP];           // 1/pivot of step k
  double pmod[GJ_MP];       // |pivot_k|^2
  int p[2];
  int sing;
};

// Pivot of column k from its values ck (one per row): the key keeps the top
// 26 bits of |re|+|im| above 64 - lane, so the maximum picks the largest
// magnitude and, among equal keys, the lowest row.  Rows already used and
// padding rows carry key 0.
__device__ __forceinline__ void gj_publish(GjShared& sh, zd ck, bool cand, int lane, int k, int buf) {
  const unsigned top = (unsigned)(__double_as_longlong(cabs1(ck)) >> 32);
  const unsigned key = cand ? ((((top >> 6) + 1u) << 6) | (unsigned)(63 - lane)) : 0u;
  const unsigned kmax = gj_wave_max(key);
  const int p = 63 - (int)(kmax & 63u);
  zd pv;
  pv.x = gj_readlane(ck.x, p);
  pv.y = gj_readlane(ck.y, p);
  sh.mbuf[buf][lane] = ck;
  if (lane == 0) {
    const double d = pv.x * pv.x + pv.y * pv.y;
    const double r = 1.0 / d;
    sh.p[buf] = p;
    sh.pinv[k] = zmk(pv.x * r, -pv.y * r);
    sh.pmod[k] = d;
    if (!(d > 0.0)) sh.sing = 1;
  }
}

__global__ __launch_bounds__(GJ_THREADS, 4) void k_gj(DevSys s, WsLayout L, double* __restrict__ ws,
                                                      double* __restrict__ Fbuf) {
  __shared__ GjShared sh;
  __shared__ double s_q[2 * GJ_MP], s_g[2 * GJ_MP];
  __shared__ double red[GJ_WAVES];
  const int w = blockIdx.x, sp = blockIdx.y;
  const int m = pick2(s.np, sp);
  if (m == 0) return;
  const int p0 = pick2(s.ipf, sp);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  double* wsw = ws + (int64_t)w * L.per_walker;
  double* dsc = wsw + pick2(L.off_det, sp);
  zd* Fg = (zd*)(Fbuf + (int64_t)w * L.f_per_walker + pick2(L.f_off, sp));
  const double* qw = wsw + L.off_q;
  for (int e = tid; e < m; e += GJ_THREADS) {
    s_q[2 * e] = qw[2 * (p0 + e)];
    s_q[2 * e + 1] = qw[2 * (p0 + e) + 1];
    s_g[2 * e] = s.gvec[2 * e];
    s_g[2 * e + 1] = s.gvec[2 * e + 1];
  }
  if (tid == 0) sh.sing = 0;
  __syncthreads();

  // A(i,j) = exp(i g_i . q_j) (paho_orbitals, reptation.f:10298-10305),
  // B_b(i,j) = i g_i^b A(i,j); rows and columns beyond m are zero
  const int i = lane;
  zd v[GJ_CW];
#pragma unroll
  for (int c = 0; c < GJ_CW; c++) {
    const int gc = c * GJ_WAVES + wave;
    const int blk = gc < m ? 0 : (gc < 2 * m ? 1 : 2);
    const int j = gc - blk * m;
    zd z = zmk(0.0, 0.0);
    if (i < m && gc < 3 * m) {
      const double kr = s_q[2 * j] * s_g[2 * i] + s_q[2 * j + 1] * s_g[2 * i + 1];
      double sv, cv;
      qsincos(kr, &sv, &cv);
      if (blk == 0) z = zmk(cv, sv);
      else {
        const double gb = s_g[2 * i + (blk - 1)];
        z = zmk(-gb * sv, gb * cv);
      }
    }
    v[c] = z;
    __builtin_amdgcn_sched_barrier(0);
  }

  bool used = i >= m;
  int myrank = -1;
  if (wave == 0) gj_publish(sh, v[0], !used, lane, 0, 0);
  __syncthreads();

  // Step k = 8 s8 + w8: slot s8 of wavefront w8 is the pivot column; slots
  // below s8 hold eliminated A columns only and are left alone.
#pragma unroll
  for (int s8 = 0; s8 < GJ_MP / GJ_WAVES; s8++) {
    for (int w8 = 0; w8 < GJ_WAVES; w8++) {
      const int k = s8 * GJ_WAVES + w8;
      if (k >= m) break;
      const int buf = k & 1;
      const int p = sh.p[buf];
      const zd pinv = sh.pinv[k];
      const zd f = sh.mbuf[buf][lane];
      zd nf = zmul(f, pinv);
      nf = zmk(-nf.x, -nf.y);
      if (lane == p) {
        used = true;
        myrank = k;
        nf = zmk(0.0, 0.0);
      }
      // the pivot row comes straight from lane p's registers into scalar
      // registers (v_readlane) and feeds the FMAs as a scalar operand; a store
      // of it to LDS by that one lane costs the LDS a full wavefront per value
#pragma unroll
      for (int c = s8; c < GJ_CW; c++) {
        const zd u = zmk(gj_readlane(v[c].x, p), gj_readlane(v[c].y, p));
        v[c] = zfma(nf, u, v[c]);
        gj_pin(v[c]);
      }
      // the owner of column k+1 publishes the next pivot
      const int kn = k + 1;
      if (kn < m && wave == (kn & (GJ_WAVES - 1))) {
        zd ck;
        if (s8 == GJ_MP / GJ_WAVES - 1) ck = v[s8];
        else ck = (w8 == GJ_WAVES - 1) ? v[s8 + 1] : v[s8];
        gj_publish(sh, ck, !used, lane, kn, kn & 1);
      }
      __syncthreads();
    }
  }
  // row i now holds piv_k times row k = myrank of X = [F_x | F_y]
  if (myrank >= 0) {
    const zd r = sh.pinv[myrank];
#pragma unroll
    for (int c = 0; c < GJ_CW; c++) {
      const int gc = c * GJ_WAVES + wave;
      if (gc < m || gc >= 3 * m) continue;
      const int blk = gc < 2 * m ? 0 : 1;
      const int j = gc - (blk + 1) * m;
      Fg[(blk * m + myrank) * m + j] = zmul(v[c], r);
    }
  }
  // log|det| = sum_k log|pivot_k| (zgedi's det, reptation.f:9412-9430)
  {
    double t = 0.0;
    for (int k = tid; k < m; k += GJ_THREADS) t += 0.5 * log(sh.pmod[k]);
    t = wave_sum(t);
    if (lane == 0) red[wave] = t;
  }
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int k = 0; k < GJ_WAVES; k++) t += red[k];
    dsc[DS_LOGDET] = t;
    dsc[DS_SING] = sh.sing ? 1.0 : 0.0;
  }
}

// Drift and diagonal Laplacian terms from F (LDS: xs[2N], dl[4m], dd[3m]).
__global__ __launch_bounds__(QMC_BLOCK) void k_fgrad(DevSys s, WsLayout L,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ ws,
                                                     const double* __restrict__ Fbuf, int tpp_log2) {
  extern __shared__ double sm[];
  __shared__ double red[QMC_BLOCK / 64];
  const int w = blockIdx.x, sp = blockIdx.y;
  const int m = pick2(s.np, sp);
  if (m == 0) return;
  const int N = s.nptot, p0 = pick2(s.ipf, sp);
  const int tid = threadIdx.x;
  double* wsw = ws + (int64_t)w * L.per_walker;
  double* dsc = wsw + pick2(L.off_det, sp);
  if (dsc[DS_SING] != 0.0) return;
  const zd* Fg = (const zd*)(Fbuf + (int64_t)w * L.f_per_walker + pick2(L.f_off, sp));
  double* xs = sm;
  double* dl = sm + 2 * N;
  double* dd = dl + 4 * m;
  const double* xw = x + (int64_t)w * 2 * N;
  for (int k = tid; k < 2 * N; k += QMC_BLOCK) xs[k] = xw[k];
  // diag(F_a F_b): four lanes per orbital j split the sum over k
  {
    const int j = tid >> 2, part = tid & 3;
    double sxx = 0, sxy = 0, syy = 0;
    if (j < m) {
      for (int k = part; k < m; k += 4) {
        const zd ax = Fg[(0 * m + j) * m + k], ay = Fg[(1 * m + j) * m + k];
        const zd bx = Fg[(0 * m + k) * m + j], by = Fg[(1 * m + k) * m + j];
        sxx += ax.x * bx.x - ax.y * bx.y;
        sxy += ax.x * by.x - ax.y * by.y;
        syy += ay.x * by.x - ay.y * by.y;
      }
    }
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
      sxx += __shfl_xor(sxx, o, 64);
      sxy += __shfl_xor(sxy, o, 64);
      syy += __shfl_xor(syy, o, 64);
    }
    if (j < m && part == 0) {
      const zd fx = Fg[(0 * m + j) * m + j], fy = Fg[(1 * m + j) * m + j];
      dl[4 * j] = fx.x;
      dl[4 * j + 1] = fx.y;
      dl[4 * j + 2] = fy.x;
      dl[4 * j + 3] = fy.y;
      dd[3 * j] = sxx;
      dd[3 * j + 1] = sxy;
      dd[3 * j + 2] = syy;
    }
  }
  __syncthreads();

  const int tpp = 1 << tpp_log2;
  const int lane = tid & (tpp - 1), grp = tid >> tpp_log2, ngrp = QMC_BLOCK >> tpp_log2;
  double acc_g2 = 0.0, acc_h0 = 0.0, acc_h1 = 0.0;
  double* gdet = dsc + DS_N;
  double* Qd = gdet + 2 * N;
  // particles in type-uniform wavefronts, partners visited type by type (as
  // in k_pairs): the table row and whether i or c belongs to this spin are
  // wavefront-uniform
  const PairSlots ps = pair_slots(s, tpp_log2);
  const int wgrp0 = __builtin_amdgcn_readfirstlane(tid >> 6) * ps.gsz;
  for (int sl0 = 0; sl0 < ps.nslot; sl0 += ngrp) {
    if (sl0 + wgrp0 >= ps.nslot) break;
    const int ti = sl0 + wgrp0 >= ps.ns0 ? 1 : 0;
    const int li = sl0 + grp - (ti ? ps.ns0 : 0);
    const bool act = li < pick2(s.np, ti);
    const int ip = pick2(s.ipf, ti) + li;
    const bool on = sp ? pick2(s.iub_flag[1], ti) : pick2(s.iub_flag[0], ti);
    const bool insp = ti == sp;
    double zxr = 0, zxi = 0, zyr = 0, zyi = 0, hq = 0, hd = 0;
    double q11 = 0, q12 = 0, q22 = 0, bx = 0, by = 0;
    if (act && on) {
      const double xi = xs[2 * ip], yi = xs[2 * ip + 1];
      for (int tj = 0; tj < s.ntypes; tj++) {
        const TabRef tb = tabref<TK_B>(s, tpair(ti, tj));
        if (tb.row0 < 0) continue;
        const bool jin = tj == sp;
        const int jb = pick2(s.ipf, tj), je = jb + pick2(s.np, tj);
        for (int c = jb + lane; c < je; c += tpp) {
          if (c == ip) continue;
          PairGeo p = pair_geo(s, xi, yi, xs[2 * c], xs[2 * c + 1]);
          double eta, deta, ddeta;
          spline3(s, tb, p.ind, p.rem, eta, deta, ddeta);
          const double cb = deta * p.byr;
          const double mb11 = eta + cb * p.ux * p.ux, mb12 = cb * p.ux * p.uy,
                       mb22 = eta + cb * p.uy * p.uy;
          const double cc = ddeta + 3.0 * cb;
          if (insp) {
            q11 += mb11; q12 += mb12; q22 += mb22;
            bx += cc * p.ux; by += cc * p.uy;
          }
          if (jin) {
            const int jf = c - p0;
            const double dxr = dl[4 * jf], dxi = dl[4 * jf + 1], dyr = dl[4 * jf + 2],
                         dyi = dl[4 * jf + 3];
            zxr -= dxr * mb11 + dyr * mb12;
            zxi -= dxi * mb11 + dyi * mb12;
            zyr -= dxr * mb12 + dyr * mb22;
            zyi -= dxi * mb12 + dyi * mb22;
            hq -= cc * (dxr * p.ux + dyr * p.uy);
            const double a11 = mb11 * mb11 + mb12 * mb12, a12 = mb12 * (mb11 + mb22),
                         a22 = mb12 * mb12 + mb22 * mb22;
            hd += dd[3 * jf] * a11 + 2.0 * dd[3 * jf + 1] * a12 + dd[3 * jf + 2] * a22;
          }
        }
      }
    }
    for (int o = 1; o < tpp; o <<= 1) {
      zxr += __shfl_xor(zxr, o, 64);
      zxi += __shfl_xor(zxi, o, 64);
      zyr += __shfl_xor(zyr, o, 64);
      zyi += __shfl_xor(zyi, o, 64);
      hq += __shfl_xor(hq, o, 64);
      hd += __shfl_xor(hd, o, 64);
      q11 += __shfl_xor(q11, o, 64);
      q12 += __shfl_xor(q12, o, 64);
      q22 += __shfl_xor(q22, o, 64);
      bx += __shfl_xor(bx, o, 64);
      by += __shfl_xor(by, o, 64);
    }
    if (act && lane == 0) {
      if (on) {
        if (insp) {
          const int jf = ip - p0;
          const double a11 = 1.0 + q11, a12 = q12, a22 = 1.0 + q22;
          const double dxr = dl[4 * jf], dxi = dl[4 * jf + 1], dyr = dl[4 * jf + 2],
                       dyi = dl[4 * jf + 3];
          zxr += dxr * a11 + dyr * a12;
          zxi += dxi * a11 + dyi * a12;
          zyr += dxr * a12 + dyr * a22;
          zyi += dxi * a12 + dyi * a22;
          hq += dxr * bx + dyr * by;
          const double b11 = a11 * a11 + a12 * a12, b12 = a12 * (a11 + a22),
                       b22 = a12 * a12 + a22 * a22;
          hd += dd[3 * jf] * b11 + 2.0 * dd[3 * jf + 1] * b12 + dd[3 * jf + 2] * b22;
          Qd[3 * jf] = q11;
          Qd[3 * jf + 1] = q12;
          Qd[3 * jf + 2] = q22;
        }
        gdet[2 * ip] = -zxr;
        gdet[2 * ip + 1] = -zyr;
        acc_g2 += pick2(s.hbs2m, ti) * (zxi * zxi + zyi * zyi);
        if (ti == 0) acc_h0 -= hq + hd;
        else acc_h1 -= hq + hd;
      } else {
        gdet[2 * ip] = 0.0;
        gdet[2 * ip + 1] = 0.0;
      }
    }
  }
  acc_g2 = block_sum(acc_g2, red);
  acc_h0 = block_sum(acc_h0, red);
  acc_h1 = block_sum(acc_h1, red);
  if (tid == 0) {
    dsc[DS_H0] = acc_h0;
    dsc[DS_H1] = acc_h1;
    dsc[DS_G2] = acc_g2;
  }
}
