// qmc_sem.hip -- single-electron VMC moves.  For trial functions without
// backflow (every backflow table zero, so quasi-coordinates are positions),
// with the two- and three-body Jastrow factors: Sherman-Morrison updates of
// the inverse Slater matrices.  With backflow: each move evaluates the
// proposal's log|psi| (further below).  The reference has no such move: its
// move displaces every electron and recomputes everything
// (reptation.f:4617-4680).  Checked against the full evaluation (the log|psi|
// change of every accepted move) and statistically against the all-electron
// walk.
//
// Moving electron j of spin s from r to r' changes column j' (its index in
// the spin) of A(k, j) = exp(i g_k . r_j) to a' and nothing else, so the
// determinant ratio is R = (A^-1 a')_j' and
// A'^-1 = A^-1 - (A^-1 a' - e_j') (row j' of A^-1) / R.  The Jastrow ratio is
// exp(-dU), dU = sum_p u(|r' - r_p|) - u(|r - r_p|) plus, with u3, the change
// of the three-body term sum_pairs (xi r)^2 - 1/2 sum_p |G_p|^2 of the full
// evaluation, G_p = sum_b xi(r_pb) (x_b - x_p), which needs only the O(N)
// changed pair terms and the stored G.  The move is accepted with
// probability |R|^2 exp(-2 dU) (a symmetric Gaussian proposal).
//
// The inverses live in HBM between calls ([nw][2][mm][mm] complex, row =
// particle, column = orbital), in FP64 or, with QMC_PREC_MIXED, in FP32 (the
// ratio and the update in FP32, the Jastrow and positions in FP64), and are
// recomputed in FP64 by qmc_sem_refresh (Gauss-Jordan with partial pivoting).
#include <cstring>

#include "qmc_internal.h"

// exp and log for the move loop, their constants kept in scalar registers
// (qk<true>) so that the loop does not hold them in vector registers next to
// the inverse.  exp: Cody-Waite reduction x = n ln2 + r (ln2 over two
// doubles), |r| <= ln2 / 2, the Taylor series of e^r through r^13
// (truncation below 5e-18), scaled by 2^n; arguments are clamped to
// [-746, 710], beyond which the result is 0 or +inf.  log: x = m 2^e with
// m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m - 1) / (m + 1),
// |s| <= 0.1716, the series through s^21 (truncation below 1e-18), plus
// e ln2; x > 0.
__device__ __forceinline__ double sem_exp(double x) {
  x = fmin(fmax(x, -746.0), 710.0);
  const double n = rint(x * qk<true>(0x1.71547652b82fep+0));
  double r = fma(-n, qk<true>(0x1.62e42fefa39efp-1), x);
  r = fma(-n, qk<true>(0x1.abc9e3b39803fp-56), r);
  double p = qk<true>(1.0 / 6227020800.0);
  p = fma(p, r, qk<true>(1.0 / 479001600.0));
  p = fma(p, r, qk<true>(1.0 / 39916800.0));
  p = fma(p, r, qk<true>(1.0 / 3628800.0));
  p = fma(p, r, qk<true>(1.0 / 362880.0));
  p = fma(p, r, qk<true>(1.0 / 40320.0));
  p = fma(p, r, qk<true>(1.0 / 5040.0));
  p = fma(p, r, qk<true>(1.0 / 720.0));
  p = fma(p, r, qk<true>(1.0 / 120.0));
  p = fma(p, r, qk<true>(1.0 / 24.0));
  p = fma(p, r, qk<true>(1.0 / 6.0));
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)n);
}

__device__ __forceinline__ double sem_log(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < qk<true>(0x1.6a09e667f3bcdp-1)) {
    m *= 2.0;
    e -= 1;
  }
  const double sv = (m - 1.0) / (m + 1.0), s2 = sv * sv;
  double p = qk<true>(1.0 / 21.0);
  p = fma(p, s2, qk<true>(1.0 / 19.0));
  p = fma(p, s2, qk<true>(1.0 / 17.0));
  p = fma(p, s2, qk<true>(1.0 / 15.0));
  p = fma(p, s2, qk<true>(1.0 / 13.0));
  p = fma(p, s2, qk<true>(1.0 / 11.0));
  p = fma(p, s2, qk<true>(1.0 / 9.0));
  p = fma(p, s2, qk<true>(1.0 / 7.0));
  p = fma(p, s2, qk<true>(1.0 / 5.0));
  p = fma(p, s2, qk<true>(1.0 / 3.0));
  const double lm = 2.0 * fma(sv * s2, p, sv);
  const double de = (double)e;
  return fma(de, qk<true>(0x1.62e42fefa39efp-1), fma(de, qk<true>(0x1.abc9e3b39803fp-56), lm));
}

template <typename Z>
struct SemZ;
template <>
struct SemZ<double2> {
  typedef double R;
};
template <>
struct SemZ<float2> {
  typedef float R;
};

// The inverse of A for (walker, spin sp): one workgroup of MR x CG threads
// per walker, A(l, k) = exp(i g_l . r_k) built in registers with the layout
// of the move kernel (thread (l, c) keeps row l at columns c, c + CG, ...),
// inverted by Gauss-Jordan exchange steps with partial pivoting, in FP64.
// Step c pivots on the row p not yet used with the largest |re| + |im| in
// column c (ties: the lowest row) and exchanges row p's output with column
// c's input: T(p, c) <- 1 / piv, T(p, k) <- -T(p, k) / piv,
// T(r, c) <- T(r, c) / piv, T(r, k) <- T(r, k) - T(r, c) T(p, k) / piv.
// After the m steps, with p_c the pivot row of step c,
// A^-1(c, p_c') = T(p_c, c'), which the threads store as they hold it.
// Rows are never moved: the pivot search is a lane-shuffle argmax per
// wavefront and one over the wavefronts in LDS, the pivot row goes through
// LDS, and each thread updates its own entries; two barriers per step.
template <typename Z, int MR, int CG>
__global__ __launch_bounds__(MR* CG) void k_sem_inverse(DevSys s, int sp, const double* __restrict__ x,
                                                        Z* __restrict__ ainv, int mm) {
  typedef typename SemZ<Z>::R R;
  constexpr int T = MR * CG, EPT = MR / CG, NWV = T / 64 > 0 ? T / 64 : 1;
  __shared__ double2 prow[MR];  // pivot row of the step
  __shared__ double cbest[NWV];
  __shared__ int cidx[NWV];
  __shared__ int pr[MR], rk[MR];  // pivot row of step c, step of pivot row l
  const int w = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = pick2(s.np, sp), p0 = pick2(s.ipf, sp);
  const int l = tid / CG, cg = tid % CG;
  const double* xw = x + ((int64_t)w * s.nptot + p0) * 2;
  const double gx = l < m ? s.gvec[2 * l] : 0.0, gy = l < m ? s.gvec[2 * l + 1] : 0.0;
  double2 b[EPT];
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int k = cg + CG * i;
    b[i] = make_double2(0.0, 0.0);
    if (l < m && k < m) {
      double sv, cv;
      qsincos(gx * xw[2 * k] + gy * xw[2 * k + 1], &sv, &cv);
      b[i] = make_double2(cv, sv);
    }
  }
  bool used = l >= m;
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    for (int c4 = 0; c4 < CG; c4++) {
      const int c = CG * i + c4;
      if (c >= m) break;
      // T(l, c), held by the thread of row l with cg == c4
      const int src = (lane & ~(CG - 1)) | c4;
      const double2 f = make_double2(__shfl(b[i].x, src, 64), __shfl(b[i].y, src, 64));
      double mag = used ? -1.0 : fabs(f.x) + fabs(f.y);
      int idx = l;
#pragma unroll
      for (int o = CG; o < 64; o <<= 1) {
        const double om = __shfl_xor(mag, o, 64);
        const int oi = __shfl_xor(idx, o, 64);
        if (om > mag || (om == mag && oi < idx)) {
          mag = om;
          idx = oi;
        }
      }
      if (lane == 0) {
        cbest[wv] = mag;
        cidx[wv] = idx;
      }
      __syncthreads();
      double bm = cbest[0];
      int p = cidx[0];
#pragma unroll
      for (int k = 1; k < NWV; k++)
        if (cbest[k] > bm || (cbest[k] == bm && cidx[k] < p)) {
          bm = cbest[k];
          p = cidx[k];
        }
      // the pivot row's threads share u_k = T(p, k) / piv (k != c) and, in
      // column c, 1 / piv; every thread then updates its entries from them
      if (l == p) {
        const double dd = 1.0 / (f.x * f.x + f.y * f.y);
        const double2 ip = make_double2(f.x * dd, -f.y * dd);
#pragma unroll
        for (int ii = 0; ii < EPT; ii++) {
          const int k = cg + CG * ii;
          const double2 q = b[ii];
          prow[k] = k == c ? ip : make_double2(q.x * ip.x - q.y * ip.y, q.x * ip.y + q.y * ip.x);
        }
        used = true;
        if (cg == 0) rk[p] = c;
      }
      if (tid == 0) pr[c] = p;
      __syncthreads();
#pragma unroll
      for (int ii = 0; ii < EPT; ii++) {
        const int k = cg + CG * ii;
        const double2 u = prow[k];
        if (l == p) b[ii] = k == c ? u : make_double2(-u.x, -u.y);
        else if (k == c) b[ii] = make_double2(f.x * u.x - f.y * u.y, f.x * u.y + f.y * u.x);
        else b[ii] = make_double2(fma(-f.x, u.x, fma(f.y, u.y, b[ii].x)), fma(-f.x, u.y, fma(-f.y, u.x, b[ii].y)));
      }
    }
  }
  __syncthreads();
  if (l < m) {
    Z* out = ainv + ((int64_t)w * 2 + sp) * mm * mm + (int64_t)rk[l] * mm;
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int k = cg + CG * i;
      if (k < m) out[pr[k]] = Z{(R)b[i].x, (R)b[i].y};
    }
  }
}

template <typename Z, int MR>
static int launch_sem_inverse(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int sp, int mm,
                              hipStream_t st) {
  hipLaunchKernelGGL((k_sem_inverse<Z, MR, 4>), dim3(w->nw), dim3(MR * 4), 0, st, c->s, sp, w->x, (Z*)d->ainv, mm);
  HIPCHK(hipGetLastError());
  return QMC_OK;
}

template <typename Z>
static int sem_inverse_spin(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int sp, int mm,
                            hipStream_t st) {
  const int m = c->s.np[sp];
  if (m <= 16) return launch_sem_inverse<Z, 16>(c, w, d, sp, mm, st);
  if (m <= 32) return launch_sem_inverse<Z, 32>(c, w, d, sp, mm, st);
  return launch_sem_inverse<Z, 64>(c, w, d, sp, mm, st);
}

// Moves of electrons j0..j1-1 of spin sp, in order: one workgroup of
// MR x CG threads per walker, the walker's m x m inverse (m <= MR) held in
// registers: thread (l, c) = (tid / CG, tid % CG) keeps row l at columns
// c, c + CG, ..., MR / CG complex entries.  The proposals come first: lane i
// of the first wavefront jumps the walker's stream 3i draws ahead and forms
// move i's Gaussian step (Box-Muller) and acceptance uniform, kept in LDS.
// Per move, threads 0..m-1 put the new orbital column a' in LDS; each thread
// forms its part of t_l = (A^-1 a')_l and the CG threads of a row sum them
// by lane shuffles; the threads of row j' put that row and R = t_j' in LDS
// while every thread sums its share of the Jastrow change; after the
// barrier every thread takes the same decision, and an accepted move
// updates each thread's entries from its own t_l and the shared row.
// Entries beyond m are zero and stay zero.  Two barriers per move, reached
// by every thread.  The log|R| of the accepted moves are summed as a
// product kept as mantissa and exponent (frexp), one log at the end.
// The FP64 inverse of a 64-row spin takes 64 registers per thread; left to
// itself the compiler spends some 110 more and fits 2 waves per SIMD.  A
// bound of 3 costs a few spilled registers and is 17% faster at N=122
// (measured, 4096 walkers).  Other shapes fit 4 or more waves unbounded (a
// bound of 5 on the FP32 64-row shape spills and is 8% slower).
template <typename Z, int MR, int CG>
struct SemWaves {
  static constexpr int value = (sizeof(Z) == 16 && MR == 64 && CG == 4) ? 3 : 1;
};

template <typename Z, int MR, int CG, bool U3>
__device__ __forceinline__ void sem_moves_walker(const DevSys& s, int w, int j0, int j1, int sp, double sigma,
                                                 double* __restrict__ x, int64_t* __restrict__ rng,
                                                 Z* __restrict__ ainv, int mm, double* __restrict__ dlogpsi,
                                                 double* __restrict__ nacc) {
  typedef typename SemZ<Z>::R R;
  constexpr int T = MR * CG, EPT = MR / CG, NWV = T / 64 > 0 ? T / 64 : 1;
  constexpr int PMAX = (128 + T - 1) / T;  // partners per thread (N <= 128)
  __shared__ Z ap[2][MR];  // new orbital column a' (two buffers: moves alternate)
  __shared__ Z rowj[MR];  // row j' of the inverse before the update
  __shared__ Z rz_sh;     // R
  __shared__ double red[3 * NWV];
  __shared__ double2 step[64];  // move i's displacement
  __shared__ double uacc[64];   // move i's acceptance uniform
  extern __shared__ __attribute__((aligned(16))) unsigned char sem_lds[];
  double* xs = reinterpret_cast<double*>(sem_lds);  // positions [N][2]
  double* Gs = xs + 2 * s.nptot;                     // U3: G_p [N][2]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int N = s.nptot, nmv = j1 - j0;
  const int m = pick2(s.np, sp), p0 = pick2(s.ipf, sp);
  const int l = tid / CG, cg = tid % CG;
  const Z* gin = ainv + ((int64_t)w * 2 + sp) * mm * mm;
  Z b[EPT];
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int k = cg + CG * i;
    b[i] = (l < m && k < m) ? gin[l * mm + k] : Z{(R)0, (R)0};
  }
  if (tid < MR) ap[0][tid] = ap[1][tid] = Z{(R)0, (R)0};
  for (int e = tid; e < 2 * N; e += T) xs[e] = x[(int64_t)w * 2 * N + e];
  const uint64_t st0 = (uint64_t)rng[w];
  if (tid < nmv) {
    // the proposal of move tid: a Gaussian step of width sigma per dimension
    uint64_t st = (st0 * lcg_pow(3 * (uint64_t)tid)) & ((1ULL << 48) - 1ULL);
    const double u1 = rannyu(st), u2 = rannyu(st), ua = rannyu(st);
    double sv, cv;
    qsincos(6.283185307179586 * u2, &sv, &cv);
    const double rad = sigma * sqrt(-2.0 * sem_log(u1));
    step[tid] = make_double2(rad * cv, rad * sv);
    uacc[tid] = ua;
  }
  double du_acc = 0.0, na = 0.0, pm = 1.0;
  int pe = 0;
  if (U3) {
    // G_p = sum_b xi(r_pb) (x_b - x_p), minimum image: the per-particle
    // vector of the three-body factor
    __syncthreads();
    for (int p = tid; p < N; p += T) {
      double gx = 0.0, gy = 0.0;
      const int tp = ptype(s, p);
      for (int q = 0; q < N; q++) {
        if (q == p) continue;
        const TabRef tu3 = tabref<TK_U3>(s, tpair(tp, ptype(s, q)));
        if (tu3.row0 < 0) continue;
        const PairGeo g = pair_geo(s, xs[2 * q], xs[2 * q + 1], xs[2 * p], xs[2 * p + 1]);
        double f, d1, d2;
        spline3(s, tu3, g.ind, g.rem, f, d1, d2);
        gx += f * g.ux;
        gy += f * g.uy;
      }
      Gs[2 * p] = gx;
      Gs[2 * p + 1] = gy;
    }
  }
  __syncthreads();
  // a'(k) = exp(i g_k . r') of a move into buffer ab, entry k by this thread
  auto put_col = [&](Z* ab, int k, double px, double py) {
    double so, co;
    qsincos_sreg(s.gvec[2 * k] * px + s.gvec[2 * k + 1] * py, &so, &co);
    ab[k] = Z{(R)co, (R)so};
  };
  // With a wavefront's worth of threads beyond the Jastrow sum's partners
  // (the last m threads, all >= N), those threads form the next move's a'
  // while the others sum, so that a move starts from a' already in LDS;
  // otherwise threads 0..m-1 form it at the start of the move.  Either way
  // from the same inputs (electron j+1 has not moved before its own move).
  // FP32 inverses only: with FP64 ones the extra live values spill inside the
  // loop, and the move measured 1% slower instead of 3% faster.
  constexpr bool PRE = sizeof(typename SemZ<Z>::R) == 4 && PMAX == 1;
  const bool pre = PRE && T - m >= N;
  for (int j = j0; j < j1; j++) {
    const int jj = j - p0, c = (j - j0) & 1;
    const double2 dx = step[j - j0];
    const double xo = xs[2 * j], yo = xs[2 * j + 1];
    const double xn = xo + dx.x, yn = yo + dx.y;
    if ((!pre || j == j0) && tid < m) put_col(ap[c], tid, xn, yn);
    __syncthreads();
    // t as explicit FMAs into two accumulator pairs (even and odd entries)
    R tr = 0, ti = 0, tr2 = 0, ti2 = 0;
#pragma unroll
    for (int i = 0; i < EPT; i += 2) {
      const Z a = ap[c][cg + CG * i];
      tr = fma(b[i].x, a.x, tr);
      tr = fma(-b[i].y, a.y, tr);
      ti = fma(b[i].x, a.y, ti);
      ti = fma(b[i].y, a.x, ti);
      if (i + 1 < EPT) {
        const Z a2 = ap[c][cg + CG * (i + 1)];
        tr2 = fma(b[i + 1].x, a2.x, tr2);
        tr2 = fma(-b[i + 1].y, a2.y, tr2);
        ti2 = fma(b[i + 1].x, a2.y, ti2);
        ti2 = fma(b[i + 1].y, a2.x, ti2);
      }
    }
    tr += tr2;
    ti += ti2;
#pragma unroll
    for (int o = 1; o < CG; o <<= 1) {
      tr += __shfl_xor(tr, o, 64);
      ti += __shfl_xor(ti, o, 64);
    }
    if (l == jj) {
#pragma unroll
      for (int i = 0; i < EPT; i++) rowj[cg + CG * i] = b[i];
      if (cg == 0) rz_sh = Z{tr, ti};
    }
    // the Jastrow change over the partners of j: the u2 terms and, with U3,
    // the three-body factor's pair terms (xi r)^2 and its -1/2 sum_p |G_p|^2
    // (G_p of each partner changes by dG_p = xi' u' - xi u, u = x_j - x_p
    // folded; G_j becomes -sum_p xi' u')
    double du = 0.0, gjx = 0.0, gjy = 0.0;
    double dgx[PMAX], dgy[PMAX];
    const double2 Gj = U3 ? ld2(Gs, j) : make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < PMAX; k++) {
      const int p = tid + k * T;
      dgx[k] = dgy[k] = 0.0;
      if (p >= N || p == j) continue;
      const int tp = tpair(sp, ptype(s, p));
      const TabRef tu2 = tabref<TK_U2>(s, tp);
      const double xp = xs[2 * p], yp = xs[2 * p + 1];
      const PairGeo pn = pair_geo(s, xn, yn, xp, yp), po = pair_geo(s, xo, yo, xp, yp);
      double fn, fo, d1, d2;
      if (tu2.row0 >= 0) {
        spline3(s, tu2, pn.ind, pn.rem, fn, d1, d2);
        spline3(s, tu2, po.ind, po.rem, fo, d1, d2);
        du += fn - fo;
      }
      if (U3) {
        const TabRef tu3 = tabref<TK_U3>(s, tp);
        if (tu3.row0 >= 0) {
          spline3(s, tu3, pn.ind, pn.rem, fn, d1, d2);
          spline3(s, tu3, po.ind, po.rem, fo, d1, d2);
          const double an = fn * pn.r, ao = fo * po.r;
          dgx[k] = fn * pn.ux - fo * po.ux;
          dgy[k] = fn * pn.uy - fo * po.uy;
          const double2 G = ld2(Gs, p);
          du += (an * an - ao * ao) - (G.x * dgx[k] + G.y * dgy[k]) - 0.5 * (dgx[k] * dgx[k] + dgy[k] * dgy[k]);
          gjx -= fn * pn.ux;
          gjy -= fn * pn.uy;
        }
      }
    }
    if (pre && j + 1 < j1 && tid >= T - m) {
      const double2 d = step[j + 1 - j0];
      put_col(ap[c ^ 1], tid - (T - m), xs[2 * (j + 1)] + d.x, xs[2 * (j + 1) + 1] + d.y);
    }
    du = wave_sum(du);
    if (U3) {
      gjx = wave_sum(gjx);
      gjy = wave_sum(gjy);
    }
    if (lane == 0) {
      red[wv] = du;
      red[NWV + wv] = gjx;
      red[2 * NWV + wv] = gjy;
    }
    __syncthreads();
    du = 0.0;
    gjx = 0.0;
    gjy = 0.0;
#pragma unroll
    for (int k = 0; k < NWV; k++) {
      du += red[k];
      gjx += red[NWV + k];
      gjy += red[2 * NWV + k];
    }
    if (U3) du -= 0.5 * ((gjx * gjx + gjy * gjy) - (Gj.x * Gj.x + Gj.y * Gj.y));
    const Z rz = rz_sh;
    const double mod2 = (double)rz.x * (double)rz.x + (double)rz.y * (double)rz.y;
    if (uacc[j - j0] < mod2 * sem_exp(-2.0 * du)) {
      // A^-1 -= (t - e_j') (row j' of A^-1) / R, 1/R as conj(R) / |R|^2
      const R inv2 = (R)(1.0 / mod2);
      const R vr = l == jj ? tr - (R)1 : tr;
      const R fr = -(vr * rz.x + ti * rz.y) * inv2, fi = -(ti * rz.x - vr * rz.y) * inv2;
#pragma unroll
      for (int i = 0; i < EPT; i++) {
        const Z q = rowj[cg + CG * i];
        b[i].x = fma(fr, q.x, b[i].x);
        b[i].x = fma(-fi, q.y, b[i].x);
        b[i].y = fma(fr, q.y, b[i].y);
        b[i].y = fma(fi, q.x, b[i].y);
      }
      if (tid == 0) {
        xs[2 * j] = xn;
        xs[2 * j + 1] = yn;
      }
      if (U3) {
#pragma unroll
        for (int k = 0; k < PMAX; k++) {
          const int p = tid + k * T;
          if (p < N && p != j) {
            Gs[2 * p] += dgx[k];
            Gs[2 * p + 1] += dgy[k];
          }
        }
        if (tid == 0) {
          Gs[2 * j] = gjx;
          Gs[2 * j + 1] = gjy;
        }
      }
      int e;
      pm = frexp(pm * mod2, &e);
      pe += e;
      du_acc += du;
      na += 1.0;
    }
  }
  __syncthreads();
  Z* gout = ainv + ((int64_t)w * 2 + sp) * mm * mm;
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int k = cg + CG * i;
    if (l < m && k < m) gout[l * mm + k] = b[i];
  }
  for (int e = tid; e < 2 * N; e += T) x[(int64_t)w * 2 * N + e] = xs[e];
  if (tid == 0) {
    rng[w] = (int64_t)((st0 * lcg_pow(3 * (uint64_t)nmv)) & ((1ULL << 48) - 1ULL));
    dlogpsi[w] += 0.5 * (sem_log(pm) + pe * 0.6931471805599453) - du_acc;
    nacc[w] += na;
  }
}

template <typename Z, int MR, int CG, bool U3>
__global__ __launch_bounds__(MR* CG, (SemWaves<Z, MR, CG>::value)) void k_sem_moves(const DevSys* __restrict__ sys,
                                                                                 int j0, int j1, int sp,
                                                                                 double sigma, double* x, int64_t* rng,
                                                                                 Z* ainv, int mm, double* dlogpsi,
                                                                                 double* nacc) {
  sem_moves_walker<Z, MR, CG, U3>(*sys, blockIdx.x, j0, j1, sp, sigma, x, rng, ainv, mm, dlogpsi, nacc);
}

// Several jobs (contexts of different N, decks and walker sets) in one
// launch: job k owns blocks w0[k] .. w0[k] + nw[k] - 1, one per walker; each
// block reads its job's system description from device memory (uniform
// loads) and runs the same per-walker body.  MR is the launch's row class
// (16, 32 or 64); a job with fewer particles leaves the extra rows idle.
struct SemJob {
  const DevSys* sys;
  double* x;
  int64_t* rng;
  void* ainv;
  double* dlogpsi;
  double* nacc;
  double sigma;
  int w0, nw, j0, j1, sp, mm;
};
#define SEM_MAX_PACK 16
struct SemJobs {
  SemJob job[SEM_MAX_PACK];
  int njobs;
};

template <typename Z, int MR, int CG, bool U3>
__global__ __launch_bounds__(MR* CG, (SemWaves<Z, MR, CG>::value)) void k_sem_moves_packed(const SemJobs jobs) {
  int k = 0;
  while (k + 1 < jobs.njobs && (int)blockIdx.x >= jobs.job[k + 1].w0) k++;
  const SemJob& jb = jobs.job[k];
  sem_moves_walker<Z, MR, CG, U3>(*jb.sys, blockIdx.x - jb.w0, jb.j0, jb.j1, jb.sp, jb.sigma, jb.x, jb.rng,
                                  (Z*)jb.ainv, jb.mm, jb.dlogpsi, jb.nacc);
}

// The context's system description in device memory, sent again before a
// launch that reads it only when it has changed since it was last sent (new
// g-vectors, time step or tuning): a blocking copy, after the work already
// queued on the stream that may still read the old one.
static int upload_sys(qmc_ctx* c, hipStream_t st) {
  if (c->sys_valid && memcmp(&c->sys_sent, &c->s, sizeof(DevSys)) == 0) return QMC_OK;
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpy(c->d_sys, &c->s, sizeof(DevSys), hipMemcpyHostToDevice));
  c->sys_sent = c->s;
  c->sys_valid = true;
  return QMC_OK;
}

template <typename Z, int MR, int CG>
static int launch_sem_moves(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int j0, int j1, int sp,
                            double sigma, int mm, hipStream_t st) {
  const bool u3 = c->s.has_u3 != 0;
  const size_t lds = (u3 ? 4 : 2) * (size_t)c->s.nptot * sizeof(double);
  if (u3)
    hipLaunchKernelGGL((k_sem_moves<Z, MR, CG, true>), dim3(w->nw), dim3(MR * CG), lds, st, c->d_sys, j0, j1, sp,
                       sigma, w->x, w->rng, (Z*)d->ainv, mm, d->dlogpsi, d->nacc);
  else
    hipLaunchKernelGGL((k_sem_moves<Z, MR, CG, false>), dim3(w->nw), dim3(MR * CG), lds, st, c->d_sys, j0, j1, sp,
                       sigma, w->x, w->rng, (Z*)d->ainv, mm, d->dlogpsi, d->nacc);
  HIPCHK(hipGetLastError());
  return QMC_OK;
}

static int sem_class(int m) { return m <= 16 ? 0 : m <= 32 ? 1 : 2; }

// The row class of the spin and the context's threads per row for it (16
// rows: 4 threads, one wavefront; 32 and 64 rows: 2 or 4).
template <typename Z>
static int sem_moves_spin(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int j0, int j1, int sp,
                          double sigma, int mm, hipStream_t st) {
  const int cls = sem_class(c->s.np[sp]);
  const int cg = c->sem_cg[sizeof(Z) == 16 ? 0 : 1][cls];
  if (cls == 0) return launch_sem_moves<Z, 16, 4>(c, w, d, j0, j1, sp, sigma, mm, st);
  if (cls == 1)
    return cg == 2 ? launch_sem_moves<Z, 32, 2>(c, w, d, j0, j1, sp, sigma, mm, st)
                   : launch_sem_moves<Z, 32, 4>(c, w, d, j0, j1, sp, sigma, mm, st);
  return cg == 2 ? launch_sem_moves<Z, 64, 2>(c, w, d, j0, j1, sp, sigma, mm, st)
                 : launch_sem_moves<Z, 64, 4>(c, w, d, j0, j1, sp, sigma, mm, st);
}

extern "C" int qmc_sem_set_tile(qmc_ctx* c, int precision, int rows, int threads_per_row) {
  if (!c) return arg_fail("null context");
  if (precision != QMC_PREC_FP64 && precision != QMC_PREC_MIXED) return arg_fail("unknown precision mode");
  const int cls = rows == 16 ? 0 : rows == 32 ? 1 : rows == 64 ? 2 : -1;
  if (cls < 0) return arg_fail("row class must be 16, 32 or 64");
  if (threads_per_row != 4 && (threads_per_row != 2 || cls == 0))
    return arg_fail("threads per row must be 4, or 2 for the 32- and 64-row classes");
  c->sem_cg[precision == QMC_PREC_MIXED ? 1 : 0][cls] = threads_per_row;
  return QMC_OK;
}

static int sem_check(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int precision) {
  if (!c || !w || !d || w->nw < 0) return arg_fail("null argument");
  if (w->nw > 0 && (!w->x || !w->rng || !d->dlogpsi || !d->nacc))
    return arg_fail("null buffer for single-electron moves");
  if (precision != QMC_PREC_FP64 && precision != QMC_PREC_MIXED) return arg_fail("unknown precision mode");
  if (!c->s.has_det) return arg_fail("single-electron moves need the Slater determinants");
  if (c->b_zero) {
    if (w->nw > 0 && !d->ainv) return arg_fail("null inverse buffer for single-electron moves");
    if (std::max(c->s.np[0], c->s.ntypes > 1 ? c->s.np[1] : 0) > QMC_MAX_ORB)
      return arg_fail("single-electron moves support at most 64 particles per spin", QMC_ERR_LIMIT);
  } else if (w->nw > 0 && !w->props) {
    return arg_fail("null walker buffer: moves with backflow keep ltf in the walkers' properties");
  } else if (!c->use_gj) {
    return arg_fail("moves with backflow need the matrix-core determinant path (QMC_NO_GJ is set)");
  }
  return QMC_OK;
}

// With backflow every quasi-coordinate depends on every electron, so one
// electron's move changes every column of both orbital matrices and no
// low-rank update applies: each move evaluates the proposed configuration's
// log|psi| (the pair kernel's Jastrow pass and quasi-coordinates, the
// determinants' pivots: no F, drift, Laplacian or potential), accepted with
// probability |psi'/psi|^2 = exp(-2 (ltf' - ltf)).  One wavefront per walker
// copies the walker with electron j displaced (the same three draws per move
// as the update path); after the evaluation one thread per walker takes the
// decision and, when accepted, moves the electron and keeps the new ltf in
// the walker's ltf slot.
__global__ __launch_bounds__(256) void k_semf_propose(DevSys s, int nw, int j, double sigma,
                                                      const double* __restrict__ x, double* __restrict__ xn,
                                                      int64_t* __restrict__ rng, double* __restrict__ ua) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + wv;
  if (w >= nw) return;
  const int N = s.nptot;
  const double* xw = x + (int64_t)w * 2 * N;
  double* xo = xn + (int64_t)w * 2 * N;
  for (int e = lane; e < 2 * N; e += 64)
    if (e != 2 * j && e != 2 * j + 1) xo[e] = xw[e];
  if (lane == 0) {
    uint64_t st = (uint64_t)rng[w];
    const double u1 = rannyu(st), u2 = rannyu(st), a = rannyu(st);
    double sv, cv;
    qsincos(6.283185307179586 * u2, &sv, &cv);
    const double rad = sigma * sqrt(-2.0 * log(u1));
    xo[2 * j] = xw[2 * j] + rad * cv;
    xo[2 * j + 1] = xw[2 * j + 1] + rad * sv;
    rng[w] = (int64_t)st;
    ua[w] = a;
  }
}

__global__ __launch_bounds__(256) void k_semf_accept(DevSys s, int nw, int j, int jltf, const double* __restrict__ xn,
                                                     const double* __restrict__ ltfn, const int32_t* __restrict__ stn,
                                                     const double* __restrict__ ua, double* __restrict__ x,
                                                     double* __restrict__ props, double* __restrict__ dlogpsi,
                                                     double* __restrict__ nacc) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  const int N = s.nptot, NP = s.n_props;
  const double lo = props[(int64_t)w * NP + jltf], ln = ltfn[w];
  // a singular proposal (status bit 2) has psi' = 0
  if ((stn[w] & 2) != 0 || !(ua[w] < exp(-2.0 * (ln - lo)))) return;
  x[(int64_t)w * 2 * N + 2 * j] = xn[(int64_t)w * 2 * N + 2 * j];
  x[(int64_t)w * 2 * N + 2 * j + 1] = xn[(int64_t)w * 2 * N + 2 * j + 1];
  props[(int64_t)w * NP + jltf] = ln;
  dlogpsi[w] += lo - ln;
  nacc[w] += 1.0;
}

static int sem_moves_full(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int j0, int j1, double sigma,
                          int precision, hipStream_t st) {
  const int nw = w->nw;
  int rc = ensure_vmc_scratch(c, nw);
  if (rc) return rc;
  for (int j = j0; j < j1; j++) {
    hipLaunchKernelGGL(k_semf_propose, dim3((nw + 3) / 4), dim3(256), 0, st, c->s, nw, j, sigma, w->x, c->d_xn,
                       w->rng, c->d_wold);
    // the proposal's ltf in d_hn (nw doubles of the proposal buffers)
    rc = launch_value(c, nw, c->d_xn, c->d_hn, c->d_stn, c->d_pn, st, precision);
    if (rc) return rc;
    hipLaunchKernelGGL(k_semf_accept, dim3((nw + 255) / 256), dim3(256), 0, st, c->s, nw, j, c->info.jltf - 1,
                       c->d_xn, c->d_hn, c->d_stn, c->d_wold, w->x, w->props, d->dlogpsi, d->nacc);
  }
  HIPCHK(hipGetLastError());
  return QMC_OK;
}

static int sem_mm(const qmc_ctx* c) { return std::max(c->s.np[0], c->s.ntypes > 1 ? c->s.np[1] : 0); }

extern "C" int qmc_sem_refresh(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int precision,
                               void* stream) {
  int rc = sem_check(c, w, d, precision);
  if (rc) return rc;
  if (w->nw == 0 || !c->b_zero) return QMC_OK;  // backflow: no inverse is carried
  HIPCHK(hipSetDevice(c->device));
  const int mm = sem_mm(c);
  for (int sp = 0; sp < c->s.ntypes; sp++) {
    if (c->s.np[sp] == 0) continue;
    rc = precision == QMC_PREC_MIXED ? sem_inverse_spin<float2>(c, w, d, sp, mm, (hipStream_t)stream)
                                     : sem_inverse_spin<double2>(c, w, d, sp, mm, (hipStream_t)stream);
    if (rc) return rc;
  }
  return QMC_OK;
}

extern "C" int qmc_sem_moves(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int j0, int j1,
                             double sigma, int precision, void* stream) {
  int rc = sem_check(c, w, d, precision);
  if (rc) return rc;
  if (j0 < 0 || j1 > c->s.nptot || j0 > j1) return arg_fail("electron range out of bounds");
  if (!(sigma > 0.0)) return arg_fail("sigma must be positive");
  if (w->nw == 0 || j0 == j1) return QMC_OK;
  HIPCHK(hipSetDevice(c->device));
  if (!c->b_zero) return sem_moves_full(c, w, d, j0, j1, sigma, precision, (hipStream_t)stream);
  rc = upload_sys(c, (hipStream_t)stream);
  if (rc) return rc;
  const int mm = sem_mm(c);
  // one launch per spin the range touches, in order
  for (int sp = 0; sp < c->s.ntypes; sp++) {
    const int a = std::max(j0, c->s.ipf[sp]), b = std::min(j1, c->s.ipf[sp] + c->s.np[sp]);
    if (a >= b) continue;
    rc = precision == QMC_PREC_MIXED ? sem_moves_spin<float2>(c, w, d, a, b, sp, sigma, mm, (hipStream_t)stream)
                                     : sem_moves_spin<double2>(c, w, d, a, b, sp, sigma, mm, (hipStream_t)stream);
    if (rc) return rc;
  }
  return QMC_OK;
}

template <typename Z, int MR, int CG, bool U3>
static int launch_sem_packed(const SemJobs& jb, int nblocks, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((k_sem_moves_packed<Z, MR, CG, U3>), dim3(nblocks), dim3(MR * CG), lds, st, jb);
  HIPCHK(hipGetLastError());
  return QMC_OK;
}

template <typename Z, int MR, int CG>
static int sem_packed_u3(const SemJobs& jb, int nblocks, bool u3, size_t lds, hipStream_t st) {
  return u3 ? launch_sem_packed<Z, MR, CG, true>(jb, nblocks, lds, st)
            : launch_sem_packed<Z, MR, CG, false>(jb, nblocks, lds, st);
}

template <typename Z>
static int sem_packed_phase(const SemJobs& jb, int nblocks, int cls, int cg, bool u3, size_t lds, hipStream_t st) {
  if (cls == 0) return sem_packed_u3<Z, 16, 4>(jb, nblocks, u3, lds, st);
  if (cls == 1)
    return cg == 2 ? sem_packed_u3<Z, 32, 2>(jb, nblocks, u3, lds, st) : sem_packed_u3<Z, 32, 4>(jb, nblocks, u3, lds, st);
  return cg == 2 ? sem_packed_u3<Z, 64, 2>(jb, nblocks, u3, lds, st) : sem_packed_u3<Z, 64, 4>(jb, nblocks, u3, lds, st);
}

extern "C" int qmc_sem_moves_packed(const qmc_sem_job* jobs, int njobs, int precision, void* stream) {
  if (!jobs || njobs < 0 || njobs > SEM_MAX_PACK) return arg_fail("between 0 and 16 jobs per packed call");
  if (njobs == 0) return QMC_OK;
  hipStream_t st = (hipStream_t)stream;
  for (int k = 0; k < njobs; k++) {
    const qmc_sem_job& jb = jobs[k];
    int rc = sem_check(jb.ctx, jb.w, jb.d, precision);
    if (rc) return rc;
    if (!jb.ctx->b_zero) return arg_fail("jobs with backflow cannot be packed (each move is a full evaluation)");
    if (jb.ctx->device != jobs[0].ctx->device) return arg_fail("packed jobs must share one device");
    if (jb.j0 < 0 || jb.j1 > jb.ctx->s.nptot || jb.j0 > jb.j1) return arg_fail("electron range out of bounds");
    if (!(jb.sigma > 0.0)) return arg_fail("sigma must be positive");
  }
  HIPCHK(hipSetDevice(jobs[0].ctx->device));
  for (int k = 0; k < njobs; k++) {
    int rc = upload_sys(jobs[k].ctx, st);
    if (rc) return rc;
  }
  // spin 0 of every job, then spin 1: each job sees its moves in the order
  // qmc_sem_moves makes them.  Within a spin, one launch per row class and
  // threads per row (as qmc_sem_moves picks for each job's context), so
  // that each job runs with the thread layout of its own call and its sums
  // are added in the same order; the three-body variant also serves jobs
  // without u3 (their u3 terms are skipped and add nothing).
  const int pi = precision == QMC_PREC_MIXED ? 1 : 0;
  for (int sp = 0; sp < 2; sp++)
  for (int cls = 0; cls < 3; cls++)
  for (int cg = 2; cg <= 4; cg += 2) {
    SemJobs jb{};
    int nblocks = 0, mr = 0;
    bool u3 = false;
    size_t lds = 0;
    for (int k = 0; k < njobs; k++) {
      qmc_ctx* c = jobs[k].ctx;
      if (sp >= c->s.ntypes || jobs[k].w->nw == 0) continue;
      const int m = c->s.np[sp];
      if (sem_class(m) != cls || (cls == 0 ? 4 : c->sem_cg[pi][cls]) != cg) continue;
      const int a = std::max(jobs[k].j0, c->s.ipf[sp]), b = std::min(jobs[k].j1, c->s.ipf[sp] + c->s.np[sp]);
      if (a >= b) continue;
      SemJob& e = jb.job[jb.njobs++];
      e.sys = c->d_sys;
      e.x = jobs[k].w->x;
      e.rng = jobs[k].w->rng;
      e.ainv = jobs[k].d->ainv;
      e.dlogpsi = jobs[k].d->dlogpsi;
      e.nacc = jobs[k].d->nacc;
      e.sigma = jobs[k].sigma;
      e.w0 = nblocks;
      e.nw = jobs[k].w->nw;
      e.j0 = a;
      e.j1 = b;
      e.sp = sp;
      e.mm = sem_mm(c);
      nblocks += e.nw;
      mr = std::max(mr, c->s.np[sp]);
      u3 = u3 || c->s.has_u3;
      lds = std::max(lds, (c->s.has_u3 ? 4 : 2) * (size_t)c->s.nptot * sizeof(double));
    }
    if (jb.njobs == 0) continue;
    (void)mr;
    const int rc = precision == QMC_PREC_MIXED ? sem_packed_phase<float2>(jb, nblocks, cls, cg, u3, lds, st)
                                               : sem_packed_phase<double2>(jb, nblocks, cls, cg, u3, lds, st);
    if (rc) return rc;
  }
  return QMC_OK;
}
