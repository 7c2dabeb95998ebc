// qmc_fgrad.hip -- drift and diagonal Laplacian terms of the backflow
// determinants for up to 64 particles per spin (reptation.f:10195-10240).
//
// The per-orbital factors d_j = F_b(j,j) and diag(A^-1 G_a G_b A)_j =
// sum_k F_a(j,k) F_b(k,j) (A^-1 G_a G_b A = F_a F_b) come from the
// determinant kernel (k_gjm stores them beside the determinant scalars);
// this kernel adds the backflow Jacobian, regenerated from the eta table,
// over the pairs.
//
// One workgroup per (walker, spin sp); every pair with a particle of sp is
// evaluated once, on the round-robin schedule of k_pairs: steps d = 1..32 of
// the spin's own block (both ends are sp particles: each takes the pair's
// part of its backflow block Q, b and of its drift and Laplacian terms
// against the other's orbital factors) and d = 0..63 across to the other
// type (the sp particle takes Q and b, the other one its drift and
// Laplacian terms).  A lane keeps its own particle's sums in registers for
// the wavefront's whole run; the partner's go to the wavefront's partial sums
// in LDS, which are added in a fixed order at the end.
#include "qmc_dev.h"

// LDS: xs[2N], dl[4][64] (F_x(j,j), F_y(j,j) as re, im), dd[3][64] (diag of
// F_x F_x, F_x F_y, F_y F_y), component-major so that lanes reading the
// factors of consecutive particles touch consecutive banks, then the
// wavefronts' partial sums where fg_off puts them (drift and Laplacian terms
// zxr, zxi, zyr, zyi, hq, hd; backflow block Q11, Q12, Q22, bx, by).
__global__ __launch_bounds__(QMC_BLOCK) void k_fgrad(DevSys s, WsLayout L,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ ws) {
  extern __shared__ __attribute__((aligned(16))) double sm[];
  __shared__ double red[QMC_BLOCK / 64];
  const int w = blockIdx.x, sp = blockIdx.y;
  if (s.nact && w >= *s.nact) return;
  const int m = pick2(s.np, sp);
  if (m == 0) return;
  const int N = s.nptot, p0 = pick2(s.ipf, sp);
  const int tid = threadIdx.x, l = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  double* wsw = ws + (int64_t)w * L.per_walker;
  double* dsc = wsw + pick2(L.off_det, sp);
  if (dsc[DS_SING] != 0.0) return;
  double* xs = sm;
  double* dl = sm + 2 * N;
  double* dd = dl + 4 * 64;
  double* parts = dd + 3 * 64;
  const double* xw = x + (int64_t)w * 2 * N;
  for (int k = tid; k < 2 * N; k += QMC_BLOCK) xs[k] = xw[k];
  const double* Fd = dsc + DS_N + 2 * N + 3 * m;
  for (int e = tid; e < m; e += QMC_BLOCK) {
#pragma unroll
    for (int c = 0; c < 4; c++) dl[c * 64 + e] = Fd[7 * e + c];
#pragma unroll
    for (int c = 0; c < 3; c++) dd[c * 64 + e] = Fd[7 * e + 4 + c];
  }
  const int ot = 1 - sp;
  const int mo = s.ntypes > 1 ? pick2(s.np, ot) : 0;
  const bool on_sp = sp ? pick2(s.iub_flag[1], sp) : pick2(s.iub_flag[0], sp);
  const bool on_ot = mo > 0 && (sp ? pick2(s.iub_flag[1], ot) : pick2(s.iub_flag[0], ot));
  const TabRef tA = tabref<TK_B>(s, tpair(sp, sp));
  const TabRef tB = tabref<TK_B>(s, tpair(sp, ot));
  // the spin's own block only feeds its own particles; the block across,
  // their Q and b and the other type's drift and Laplacian terms
  const int nA = (on_sp && tA.row0 >= 0) ? 32 : 0;
  const int nB = (mo > 0 && tB.row0 >= 0 && (on_sp || on_ot)) ? 64 : 0;
  for (int k = tid; k < fg_off(nA, nB, 4); k += QMC_BLOCK) parts[k] = 0.0;
  __syncthreads();
  const int u0 = fg_u(nA, nB, wv), u1 = fg_u(nA, nB, wv + 1);
  const bool hasA = fg_hasA(nA, nB, wv);
  double* pq_w = parts + fg_off(nA, nB, wv);                     // qb[FG_NQ][64]
  double* zs_w = pq_w + FG_NQ * 64;                              // zs[FG_NZ][64]
  double* zo_w = zs_w + (hasA ? FG_NZ * 64 : 0);                 // zo[FG_NZ][64]
  // lane l's particle of the spin: its position and orbital factors
  const bool li = l < m;
  const int ip = p0 + l;
  const double xi = li ? xs[2 * ip] : 0.0, yi = li ? xs[2 * ip + 1] : 0.0;
  double dlr[4] = {0, 0, 0, 0}, ddr[3] = {0, 0, 0};
  if (li) {
#pragma unroll
    for (int c = 0; c < 4; c++) dlr[c] = dl[c * 64 + l];
#pragma unroll
    for (int c = 0; c < 3; c++) ddr[c] = dd[c * 64 + l];
  }
  double iz[FG_NZ] = {0, 0, 0, 0, 0, 0}, iq[FG_NQ] = {0, 0, 0, 0, 0};
  for (int u = u0; u < u1; u++) {
    const bool blkA = u < nA;
    const int d = blkA ? u + 1 : u - nA;
    const int jl = (l + d) & 63;
    const bool ok = li && (blkA ? (jl < m && (d < 32 || l < 32)) : jl < mo);
    if (!ok) continue;
    const int jp = blkA ? p0 + jl : pick2(s.ipf, ot) + jl;
    double* pz = (blkA ? zs_w : zo_w) + jl;
    // r = x_i - x_j; the pair's Jacobian block mb (even in r), cc r (odd)
    const double2 xj2 = ld2(xs, jp);
    const PairGeo p = pair_geo(s, xi, yi, xj2.x, xj2.y);
    double eta, deta, ddeta;
    spline3(s, blkA ? tA : tB, p.ind, p.rem, eta, deta, ddeta);
    const double cb = deta * p.byr;
    const double mb11 = eta + cb * p.ux * p.ux, mb12 = cb * p.ux * p.uy, mb22 = eta + cb * p.uy * p.uy;
    const double cc = ddeta + 3.0 * cb;
    const double a11 = mb11 * mb11 + mb12 * mb12, a12 = mb12 * (mb11 + mb22), a22 = mb12 * mb12 + mb22 * mb22;
    // i's backflow block
    iq[0] += mb11;
    iq[1] += mb12;
    iq[2] += mb22;
    iq[3] += cc * p.ux;
    iq[4] += cc * p.uy;
    // j's drift and Laplacian terms against i's orbital factors (r_ji = -r)
    pz[0 * 64] -= dlr[0] * mb11 + dlr[2] * mb12;
    pz[1 * 64] -= dlr[1] * mb11 + dlr[3] * mb12;
    pz[2 * 64] -= dlr[0] * mb12 + dlr[2] * mb22;
    pz[3 * 64] -= dlr[1] * mb12 + dlr[3] * mb22;
    pz[4 * 64] += cc * (dlr[0] * p.ux + dlr[2] * p.uy);
    pz[5 * 64] += ddr[0] * a11 + 2.0 * ddr[1] * a12 + ddr[2] * a22;
    if (blkA) {
      // j is a particle of the spin too: its backflow block, and i's terms
      // against j's orbital factors
      double* pq = pq_w + jl;
      pq[0 * 64] += mb11;
      pq[1 * 64] += mb12;
      pq[2 * 64] += mb22;
      pq[3 * 64] -= cc * p.ux;
      pq[4 * 64] -= cc * p.uy;
      const double dxr = dl[jl], dxi = dl[64 + jl], dyr = dl[128 + jl], dyi = dl[192 + jl];
      iz[0] -= dxr * mb11 + dyr * mb12;
      iz[1] -= dxi * mb11 + dyi * mb12;
      iz[2] -= dxr * mb12 + dyr * mb22;
      iz[3] -= dxi * mb12 + dyi * mb22;
      iz[4] -= cc * (dxr * p.ux + dyr * p.uy);
      iz[5] += dd[jl] * a11 + 2.0 * dd[64 + jl] * a12 + dd[128 + jl] * a22;
    }
  }
  if (li && u1 > u0) {
    if (hasA) {
#pragma unroll
      for (int q = 0; q < FG_NZ; q++) zs_w[q * 64 + l] += iz[q];
    }
#pragma unroll
    for (int q = 0; q < FG_NQ; q++) pq_w[q * 64 + l] += iq[q];
  }
  __syncthreads();

  double acc_g2 = 0.0, acc_h0 = 0.0, acc_h1 = 0.0;
  double* gdet = dsc + DS_N;
  double* Qd = gdet + 2 * N;
  for (int pp = tid; pp < N; pp += QMC_BLOCK) {
    const int ti = ptype(s, pp);
    const bool insp = ti == sp;
    if (!insp && (s.ntypes < 2 || mo == 0)) continue;
    const bool on = insp ? on_sp : on_ot;
    if (!on) {
      gdet[2 * pp] = 0.0;
      gdet[2 * pp + 1] = 0.0;
      continue;
    }
    // the wavefronts' partial sums, added in wavefront order
    const int jx = insp ? pp - p0 : pp - pick2(s.ipf, ot);
    double z[FG_NZ] = {0, 0, 0, 0, 0, 0}, q[FG_NQ] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 4; v++) {
      if (fg_u(nA, nB, v + 1) == fg_u(nA, nB, v)) continue;
      const double* pv = parts + fg_off(nA, nB, v);
      const bool hA = fg_hasA(nA, nB, v);
      if (insp) {
#pragma unroll
        for (int c = 0; c < FG_NQ; c++) q[c] += pv[c * 64 + jx];
      }
      if (insp ? hA : fg_hasB(nA, nB, v)) {
        const double* zv = pv + FG_NQ * 64 + (insp ? 0 : (hA ? FG_NZ * 64 : 0));
#pragma unroll
        for (int c = 0; c < FG_NZ; c++) z[c] += zv[c * 64 + jx];
      }
    }
    if (insp) {
      const int jf = jx;
      const double a11 = 1.0 + q[0], a12 = q[1], a22 = 1.0 + q[2];
      const double dxr = dl[jf], dxi = dl[64 + jf], dyr = dl[128 + jf], dyi = dl[192 + jf];
      z[0] += dxr * a11 + dyr * a12;
      z[1] += dxi * a11 + dyi * a12;
      z[2] += dxr * a12 + dyr * a22;
      z[3] += dxi * a12 + dyi * a22;
      z[4] += dxr * q[3] + dyr * q[4];
      const double b11 = a11 * a11 + a12 * a12, b12 = a12 * (a11 + a22), b22 = a12 * a12 + a22 * a22;
      z[5] += dd[jf] * b11 + 2.0 * dd[64 + jf] * b12 + dd[128 + jf] * b22;
      Qd[3 * jf] = q[0];
      Qd[3 * jf + 1] = q[1];
      Qd[3 * jf + 2] = q[2];
    }
    gdet[2 * pp] = -z[0];
    gdet[2 * pp + 1] = -z[2];
    acc_g2 += pick2(s.hbs2m, ti) * (z[1] * z[1] + z[3] * z[3]);
    if (ti == 0) acc_h0 -= z[4] + z[5];
    else acc_h1 -= z[4] + z[5];
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
