// qmc_gjm.hip -- determinant path for up to 64 particles per spin on the
// FP64 matrix cores (the default at every size; k_det is the LDS LU
// alternative).
//
// k_gjm computes F_b = A^-1 (G_b A) (zv of reptation.f:10182-10194, with
// A(i,j) = exp(i g_i . q_j), G_b = diag(i g^b)) and log|det A| in two phases,
// both built from v_mfma_f64_16x16x4f64 tiles (real arithmetic: a complex
// product is four real MFMAs).  One workgroup of 8 wavefronts per
// (walker, spin); the 64 x 64 matrix (rows = orbitals, columns = particles,
// zero beyond m) is split into 4 x 4 complex 16 x 16 tiles, wavefront w
// holding column tile w % 4 of row tiles 2 (w / 4) and 2 (w / 4) + 1.
//
// Phase 1, in-place Gauss-Jordan inverse in blocks of 4 columns.  Pivot
// choice is zgefa's (largest |re|+|im| among the rows not yet used,
// reptation.f:9286-9298; rows whose magnitudes agree to 13 bits may be taken
// in another order, which changes rounding only).  For the block of columns
// k0..k0+3 one wavefront runs the 4 pivot steps on the 64 x 4 panel (lane =
// row) and accumulates the block's row operations as T = I + F' E^T, E
// selecting the 4 pivot rows: row r of the matrix becomes
// row r + sum_j F'(r,j) (pivot row p_j as it was before the block).  The
// whole block is then one rank-4 update M += F' U on the matrix cores, U
// holding the 4 pivot rows; the panel columns, whose A values are no longer
// needed, take the columns e_{p_j} of the identity instead (U = identity
// there), so that after the last block column k of the matrix holds column
// p_k of the accumulated transform S (S A has row p_k equal to piv_k e_k).
// The tiles live in the accumulators; after every block each wavefront
// also stores them to an LDS copy of the matrix, from which the next
// block's panel and pivot rows are read, so a block costs two barriers.
//
// Phase 2: A^-1 = D^-1 P S with P the pivot permutation, so
// F_b(k, .) = (S G'_b)(p_k, .) / piv_k with G'_b(k', j) = i g^b_{p_k'} A(p_k', j):
// two complex 64 x 64 x 64 products on the matrix cores, S taken from the
// LDS copy and G'_b regenerated from q and g.  log|det A| = sum_k log|piv_k|
// (zgedi's det, reptation.f:9412-9430).  The per-orbital terms of the drift
// and diagonal Laplacian (reptation.f:10195-10240) need F_b(j,j) and
// diag(F_a F_b) = diag(A^-1 G_a G_b A); both are formed here and stored
// beside the determinant scalars (7 doubles per orbital), so the next kernel
// does not read F.
#include "qmc_dev.h"

typedef double gm_d4 __attribute__((ext_vector_type(4)));

#define GM_THREADS 512
#define GM_PS 5  // row stride (complex) of the F' buffer

struct GmShared {
  double q[128], g[128];    // quasi-coordinates of the spin's particles, g-vectors
  zd pinv[64];              // 1 / piv_k
  double pmod[64];          // |piv_k|^2
  int prow[64], rank[64];   // pivot row of step k, step of pivot row r
  int sing;
  int used[64];             // row r already a pivot row
  // the matrix, real and imaginary planes: (r, c) at r * 64 + (c ^ gm_swz(r))
  double Mre[64 * 64], Mim[64 * 64];
  zd Fl[64 * GM_PS];        // F' of the block, row r at r * GM_PS
  double Ure[4 * 64], Uim[4 * 64];  // the block's pivot rows before the block
};

// Column swizzle of row r.  With it the accesses below touch distinct banks:
// tile stores (2 consecutive rows x 16 columns), MFMA operand loads (16 rows
// x 2 columns), panel loads (32 consecutive rows, one column) and row copies.
__device__ __forceinline__ int gm_swz(int r) { return (r & 15) | (((r ^ (r >> 4)) & 1) << 4); }

__device__ __forceinline__ gm_d4 gm_mfma(double a, double b, gm_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 1/d for d > 0: hardware reciprocal and two Newton steps.
__device__ __forceinline__ double gm_rcp(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

// Store the wavefront's tiles into the LDS copy of the matrix.
__device__ __forceinline__ void gm_put_tiles(GmShared& sh, const gm_d4 (&acc)[2][2], int rh, int col, int lr) {
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = 16 * (2 * rh + t) + lr + 4 * v;
      const int a = r * 64 + (col ^ gm_swz(r));
      sh.Mre[a] = acc[t][0][v];
      sh.Mim[a] = acc[t][1][v];
    }
}

// The pivot steps of the block of columns k0..k0+nb-1 on its panel P (lane =
// row, every lane of the wavefront active): pivot rows into pj and the LDS
// pivot tables, the block's transform into Fp (F'(r, j) in lane r), and the
// pivot rows as they were before the block into Ur (lane = column), loaded
// from the LDS copy as soon as each pivot is known.
__device__ __forceinline__ void gm_pivot_steps(GmShared& sh, zd (&P)[4], zd (&Fp)[4], zd (&Ur)[4], int (&pj)[4],
                                               bool& used, int k0, int nb, int lane) {
#pragma unroll
  for (int j = 0; j < 4; j++) Fp[j] = zmk(0.0, 0.0);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (j < nb) {
      // pivot row: largest |re|+|im| among unused rows, lowest row on a tie
      const unsigned top = (unsigned)(__double_as_longlong(cabs1(P[j])) >> 32);
      const unsigned key = used ? 0u : ((((top >> 6) + 1u) << 6) | (unsigned)(63 - lane));
      const int p = 63 - (int)(gj_wave_max(key) & 63u);
      pj[j] = p;
      {
        const int a = p * 64 + (lane ^ gm_swz(p));
        Ur[j] = zmk(sh.Mre[a], sh.Mim[a]);
      }
      const zd pv = zmk(gj_readlane(P[j].x, p), gj_readlane(P[j].y, p));
      const double d = pv.x * pv.x + pv.y * pv.y;
      const double rd = gm_rcp(d);
      const zd pinv = zmk(pv.x * rd, -pv.y * rd);
      zd f = zmul(P[j], pinv);
      if (lane == p) f = zmk(0.0, 0.0);
      const zd nf = zmk(-f.x, -f.y);
#pragma unroll
      for (int jj = j + 1; jj < 4; jj++) {
        const zd u = zmk(gj_readlane(P[jj].x, p), gj_readlane(P[jj].y, p));
        P[jj] = zfma(nf, u, P[jj]);
      }
#pragma unroll
      for (int i = 0; i < j; i++) {
        const zd u = zmk(gj_readlane(Fp[i].x, p), gj_readlane(Fp[i].y, p));
        Fp[i] = zfma(nf, u, Fp[i]);
      }
      Fp[j] = nf;
      if (lane == p) used = true;
      if (lane == 0) {
        sh.prow[k0 + j] = p;
        sh.rank[p] = k0 + j;
        sh.pinv[k0 + j] = pinv;
        sh.pmod[k0 + j] = d;
        if (!(d > 0.0)) sh.sing = 1;
      }
    }
  }
}

template <bool PART, bool VAL>
__device__ __forceinline__ void gjm_run(DevSys s, WsLayout L, double* __restrict__ ws,
                                        double* __restrict__ Fbuf) {
  extern __shared__ __attribute__((aligned(16))) unsigned char gm_lds[];
  GmShared& sh = *reinterpret_cast<GmShared*>(gm_lds);
  const int w = blockIdx.x, sp = blockIdx.y;
  if (s.nact && w >= *s.nact) return;
  const int m = pick2(s.np, sp);
  if (m == 0) return;
  const int p0 = pick2(s.ipf, sp);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ct = wave & 3, rh = wave >> 2;
  // MFMA lane roles: accumulator rows lr + 4v, column lc; operand k index lr
  const int lr = lane >> 4, lc = lane & 15;
  const int col = 16 * ct + lc;
  // PART: tiles wholly beyond row or column m start at zero and stay zero
  // (the row operations of rows >= m and the pivot rows' entries in columns
  // >= m are zero), and nothing is read from them for rows or columns < m:
  // their products are skipped.  live2: both of the wavefront's row tiles
  // take part, live1: the first only.
  const int ntile = (m + 15) >> 4;
  const bool live1 = !PART || (ct < ntile && 2 * rh < ntile);
  const bool live2 = !PART || (live1 && 2 * rh + 1 < ntile);
  const double* qw = ws + (int64_t)w * L.per_walker + L.off_q;
  for (int e = tid; e < 64; e += GM_THREADS) {
    const bool in = e < m;
    sh.q[2 * e] = in ? qw[2 * (p0 + e)] : 0.0;
    sh.q[2 * e + 1] = in ? qw[2 * (p0 + e) + 1] : 0.0;
    sh.g[2 * e] = in ? s.gvec[2 * e] : 0.0;
    sh.g[2 * e + 1] = in ? s.gvec[2 * e + 1] : 0.0;
  }
  if (tid == 0) sh.sing = 0;
  __syncthreads();

  // A(i,j) = exp(i g_i . q_j) (paho_orbitals, reptation.f:10298-10305); [t][0]
  // real and [t][1] imaginary parts of row tile 2 rh + t, column tile ct
  gm_d4 acc[2][2];
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = 16 * (2 * rh + t) + lr + 4 * v;
      double sv = 0.0, cv = 0.0;
      if (r < m && col < m) qsincos(sh.q[2 * col] * sh.g[2 * r] + sh.q[2 * col + 1] * sh.g[2 * r + 1], &sv, &cv);
      acc[t][0][v] = cv;
      acc[t][1][v] = sv;
    }

  // ---- phase 1.  The pivot steps of block k0 run on wavefront (k0 / 4) mod 4,
  // so that they fall on each SIMD in turn.
  if (tid < 64) sh.used[tid] = tid >= m ? 1 : 0;
  gm_put_tiles(sh, acc, rh, col, lr);
  __syncthreads();
  for (int k0 = 0; k0 < m; k0 += 4) {
    const int nb = min(4, m - k0);
    if (wave == ((k0 >> 2) & 3)) {
      // the pivot steps are the chain every wavefront of the workgroup waits
      // on: they issue ahead of the other workgroup's products on this SIMD
      __builtin_amdgcn_s_setprio(1);
      const int sw = gm_swz(lane);
      zd P[4], Fp[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int a = lane * 64 + ((k0 + j) ^ sw);
        P[j] = zmk(sh.Mre[a], sh.Mim[a]);
      }
      bool used = sh.used[lane] != 0;
      int pj[4] = {0, 0, 0, 0};
      zd Ur[4];
      gm_pivot_steps(sh, P, Fp, Ur, pj, used, k0, nb, lane);
      sh.used[lane] = used ? 1 : 0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (j < nb) {
          sh.Ure[j * 64 + lane] = Ur[j].x;
          sh.Uim[j * 64 + lane] = Ur[j].y;
        }
      }
#pragma unroll
      for (int j = 0; j < 4; j++) sh.Fl[lane * GM_PS + j] = Fp[j];
      __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();
    // the block's 4 pivot rows in one LDS read (entries beyond nb unused)
    const int4 pr4 = *reinterpret_cast<const int4*>(&sh.prow[k0]);
    const int pb[4] = {__builtin_amdgcn_readfirstlane(pr4.x), __builtin_amdgcn_readfirstlane(pr4.y),
                       __builtin_amdgcn_readfirstlane(pr4.z), __builtin_amdgcn_readfirstlane(pr4.w)};
    // the panel columns become e_{p_j}
    const bool panel_tile = ct == (k0 >> 4);
    const int jj = lc - (k0 & 15);
    if (panel_tile && jj >= 0 && jj < nb) {
      const int pr = jj == 0 ? pb[0] : (jj == 1 ? pb[1] : (jj == 2 ? pb[2] : pb[3]));
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = 16 * (2 * rh + t) + lr + 4 * v;
          acc[t][0][v] = r == pr ? 1.0 : 0.0;
          acc[t][1][v] = 0.0;
        }
    }
    // M += F' U: B operand U(lr, col) (identity in the panel columns), A
    // operand F'(row, lr)
    zd uo = lr < nb ? zmk(sh.Ure[lr * 64 + col], sh.Uim[lr * 64 + col]) : zmk(0.0, 0.0);
    if (panel_tile && jj >= 0 && jj < 4) uo = zmk(lr == jj && jj < nb ? 1.0 : 0.0, 0.0);
    zd fo[2];
#pragma unroll
    for (int t = 0; t < 2; t++) fo[t] = sh.Fl[(16 * (2 * rh + t) + lc) * GM_PS + lr];
    auto upd = [&](auto ntc) {
      constexpr int NT = decltype(ntc)::value;
#pragma unroll
      for (int t = 0; t < NT; t++) {
        acc[t][0] = gm_mfma(fo[t].x, uo.x, acc[t][0]);
        acc[t][1] = gm_mfma(fo[t].x, uo.y, acc[t][1]);
      }
#pragma unroll
      for (int t = 0; t < NT; t++) {
        acc[t][0] = gm_mfma(-fo[t].y, uo.y, acc[t][0]);
        acc[t][1] = gm_mfma(fo[t].y, uo.x, acc[t][1]);
      }
    };
    if (live2) upd(std::integral_constant<int, 2>{});
    else if (live1) upd(std::integral_constant<int, 1>{});
    gm_put_tiles(sh, acc, rh, col, lr);
    __syncthreads();
  }

  // ---- phase 2 (not for a value-only evaluation, which needs log|det A| alone)
  if (!VAL) {
    const int j = col;
    const double qx = sh.q[2 * j], qy = sh.q[2 * j + 1];
    double* Fd = ws + (int64_t)w * L.per_walker + pick2(L.off_det, sp) + DS_N + 2 * s.nptot + 3 * m;
    // H_b = S G'_b for output column tile ct (particles j), row tiles 2 rh + t:
    // h[t][b][0] real, h[t][b][1] imaginary.  Beside it the wavefronts of row
    // half 0 form diag(F_a F_b)_j = diag(A^-1 G_a G_b A)_j
    // = sum_k' S(p_j, k') (-g^a g^b)_{p_k'} A(p_k', j) / piv_j from the same
    // A(p_k', j) values (lane (lr, lc) summing k' = lr mod 4 for particle j, the
    // four partial sums combined after the loop), O(m^2) in all.  The sin/cos
    // constants stay in scalar registers here, where the accumulators take the
    // vector ones.
    const bool dg = rh == 0;
    int prj = 0, swj = 0;
    zd pij = zmk(0.0, 0.0);
    if (dg && j < m) {
      prj = sh.prow[j];
      swj = gm_swz(prj);
      pij = sh.pinv[j];
    }
    double dxx = 0.0, dxy = 0.0, dyy = 0.0;
    gm_d4 h[2][2][2];
  #pragma unroll
    for (int t = 0; t < 2; t++)
  #pragma unroll
      for (int b = 0; b < 2; b++) h[t][b][0] = h[t][b][1] = gm_d4{0.0, 0.0, 0.0, 0.0};
    // a wavefront with no live tile has neither outputs nor diagonal terms
    // (its particles are >= m)
    for (int ks = 0; live1 && 4 * ks < m; ks++) {
      const int kp = 4 * ks + lr;
      double bx_re = 0.0, bx_im = 0.0, by_re = 0.0, by_im = 0.0;
      if (kp < m && j < m) {
        const int o = sh.prow[kp];
        const double gx = sh.g[2 * o], gy = sh.g[2 * o + 1];
        double sv, cv;
        qsincos_sreg(gx * qx + gy * qy, &sv, &cv);
        bx_re = -gx * sv;
        bx_im = gx * cv;
        by_re = -gy * sv;
        by_im = gy * cv;
        if (dg) {
          const int ad = prj * 64 + (kp ^ swj);
          const double wr = zmul(zmul(zmk(sh.Mre[ad], sh.Mim[ad]), pij), zmk(cv, sv)).x;
          dxx = fma(-gx * gx, wr, dxx);
          dxy = fma(-gx * gy, wr, dxy);
          dyy = fma(-gy * gy, wr, dyy);
        }
      }
      zd a[2];
  #pragma unroll
      for (int t = 0; t < 2; t++) {
        const int r = 16 * (2 * rh + t) + lc;
        const int ad = r * 64 + (kp ^ gm_swz(r));
        a[t] = zmk(sh.Mre[ad], sh.Mim[ad]);
      }
  #pragma unroll
      for (int t = 0; t < 2; t++) {
        h[t][0][0] = gm_mfma(a[t].x, bx_re, h[t][0][0]);
        h[t][0][1] = gm_mfma(a[t].x, bx_im, h[t][0][1]);
        h[t][1][0] = gm_mfma(a[t].x, by_re, h[t][1][0]);
        h[t][1][1] = gm_mfma(a[t].x, by_im, h[t][1][1]);
      }
  #pragma unroll
      for (int t = 0; t < 2; t++) {
        h[t][0][0] = gm_mfma(-a[t].y, bx_im, h[t][0][0]);
        h[t][0][1] = gm_mfma(a[t].y, bx_re, h[t][0][1]);
        h[t][1][0] = gm_mfma(-a[t].y, by_im, h[t][1][0]);
        h[t][1][1] = gm_mfma(a[t].y, by_re, h[t][1][1]);
      }
    }
    if (dg) {
  #pragma unroll
      for (int o = 16; o < 64; o <<= 1) {
        dxx += __shfl_xor(dxx, o, 64);
        dxy += __shfl_xor(dxy, o, 64);
        dyy += __shfl_xor(dyy, o, 64);
      }
      if (j < m && lr == 0) {
        Fd[7 * j + 4] = dxx;
        Fd[7 * j + 5] = dxy;
        Fd[7 * j + 6] = dyy;
      }
    }
    // F_b(k, j) = H_b(p_k, j) / piv_k
    zd* Fg = (zd*)(Fbuf + (int64_t)w * L.f_per_walker + pick2(L.f_off, sp));
    if (j < m) {
  #pragma unroll
      for (int t = 0; t < 2; t++)
  #pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = 16 * (2 * rh + t) + lr + 4 * v;
          if (r < m) {
            const int k = sh.rank[r];
            const zd pi = sh.pinv[k];
  #pragma unroll
            for (int b = 0; b < 2; b++) {
              const zd f = zmul(zmk(h[t][b][0][v], h[t][b][1][v]), pi);
              Fg[(b * m + k) * m + j] = f;
              // the diagonal F_b(j, j) also goes to the per-orbital terms
              if (k == j) {
                Fd[7 * j + 2 * b] = f.x;
                Fd[7 * j + 2 * b + 1] = f.y;
              }
            }
          }
        }
    }
  }
  if (wave == 0) {
    double* dsc = ws + (int64_t)w * L.per_walker + pick2(L.off_det, sp);
    double t = 0.0;
    for (int k = lane; k < m; k += 64) t += 0.5 * log(sh.pmod[k]);
    t = wave_sum(t);
    if (lane == 0) {
      dsc[DS_LOGDET] = t;
      dsc[DS_SING] = sh.sing ? 1.0 : 0.0;
    }
  }
}

// every spin 49..64 particles: all 4 x 4 tiles take part
__global__ __launch_bounds__(GM_THREADS, 4) void k_gjm(DevSys s, WsLayout L, double* __restrict__ ws,
                                                       double* __restrict__ Fbuf) {
  gjm_run<false, false>(s, L, ws, Fbuf);
}

// a spin of 1..48 particles: the products of the tiles beyond m are skipped
__global__ __launch_bounds__(GM_THREADS, 4) void k_gjm_part(DevSys s, WsLayout L,
                                                            double* __restrict__ ws,
                                                            double* __restrict__ Fbuf) {
  gjm_run<true, false>(s, L, ws, Fbuf);
}

// phase 1 and log|det A| only (the value of the wavefunction)
__global__ __launch_bounds__(GM_THREADS, 4) void k_gjm_val(DevSys s, WsLayout L, double* __restrict__ ws,
                                                           double* __restrict__ Fbuf) {
  gjm_run<false, true>(s, L, ws, Fbuf);
}

__global__ __launch_bounds__(GM_THREADS, 4) void k_gjm_part_val(DevSys s, WsLayout L,
                                                                double* __restrict__ ws,
                                                                double* __restrict__ Fbuf) {
  gjm_run<true, true>(s, L, ws, Fbuf);
}

size_t gjm_lds_bytes() { return sizeof(GmShared); }
