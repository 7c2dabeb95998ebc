// qmc_fgrad.hip -- drift and diagonal Laplacian terms of the backflow
// determinants for 33..64 particles per spin (reptation.f:10195-10240).
//
// The per-orbital factors d_j = F_b(j,j) and diag(A^-1 G_a G_b A)_j =
// sum_k F_a(j,k) F_b(k,j) (A^-1 G_a G_b A = F_a F_b) come from the
// determinant kernel (k_gjm stores them beside the determinant scalars);
// this kernel adds the backflow Jacobian, regenerated from the eta table,
// over the pairs.
#include "qmc_dev.h"

// LDS: xs[2N], dl[4m] (F_x(j,j), F_y(j,j)), dd[3m] (diag of F_x F_x, F_x F_y, F_y F_y).
__global__ __launch_bounds__(QMC_BLOCK) void k_fgrad(DevSys s, WsLayout L,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ ws, int tpp_log2) {
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
  double* xs = sm;
  double* dl = sm + 2 * N;
  double* dd = dl + 4 * m;
  const double* xw = x + (int64_t)w * 2 * N;
  for (int k = tid; k < 2 * N; k += QMC_BLOCK) xs[k] = xw[k];
  const double* Fd = dsc + DS_N + 2 * N + 3 * m;
  for (int e = tid; e < m; e += QMC_BLOCK) {
#pragma unroll
    for (int c = 0; c < 4; c++) dl[4 * e + c] = Fd[7 * e + c];
#pragma unroll
    for (int c = 0; c < 3; c++) dd[3 * e + c] = Fd[7 * e + 4 + c];
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
        const bool jin = tj == sp;
        // pairs with neither particle in this spin contribute nothing
        if (tb.row0 < 0 || !(insp || jin)) continue;
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
