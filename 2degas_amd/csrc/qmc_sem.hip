// qmc_sem.hip -- single-electron VMC moves with Sherman-Morrison updates of
// the inverse Slater matrices, for trial functions without backflow (every
// backflow table zero, so quasi-coordinates are positions) and with the
// two-body Jastrow u2.  The reference has no such move: its move displaces
// every electron and recomputes everything (reptation.f:4617-4680); this is
// the single-electron scheme of the plain Slater-Jastrow function, checked
// against the full evaluation (the log|psi| change of every accepted move)
// and statistically against the all-electron walk.
//
// Moving electron j of spin s from r to r' changes column j' (its index in
// the spin) of A(k, j) = exp(i g_k . r_j) to a' and nothing else, so the
// determinant ratio is R = (A^-1 a')_j' and
// A'^-1 = A^-1 - (A^-1 a' - e_j') (row j' of A^-1) / R.  The Jastrow ratio is
// exp(-dU), dU = sum_p u(|r' - r_p|) - u(|r - r_p|).  The move is accepted
// with probability |R|^2 exp(-2 dU) (a symmetric Gaussian proposal).
//
// The inverses live in HBM between calls ([nw][2][mm][mm] complex, row =
// particle, column = orbital), in FP64 or, with QMC_PREC_MIXED, in FP32 (the
// ratio and the update in FP32, the Jastrow and positions in FP64), and are
// recomputed in FP64 by qmc_sem_refresh (Gauss-Jordan with partial pivoting).
#include "qmc_internal.h"

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

// The inverse of A for (walker, spin): A built from the positions in LDS
// and inverted in place by Gauss-Jordan with partial pivoting (largest
// |re| + |im| in the column), the column order restored at the end.
template <typename Z>
__global__ __launch_bounds__(256) void k_sem_inverse(DevSys s, int nw, const double* __restrict__ x,
                                                     Z* __restrict__ ainv, int mm) {
  extern __shared__ __attribute__((aligned(16))) unsigned char sem_lds[];
  double2* A = reinterpret_cast<double2*>(sem_lds);  // m x m, (r, c) at r * m + c
  __shared__ int perm[64];
  __shared__ double2 colc[64];
  __shared__ int sh_piv;
  const int w = blockIdx.x, sp = blockIdx.y, tid = threadIdx.x;
  if (w >= nw) return;
  const int m = pick2(s.np, sp);
  if (m == 0) return;
  const int p0 = pick2(s.ipf, sp);
  const double* xw = x + (int64_t)w * 2 * s.nptot;
  // A(k, j) = exp(i g_k . r_j)
  for (int e = tid; e < m * m; e += 256) {
    const int k = e / m, j = e % m;
    double sv, cv;
    qsincos(s.gvec[2 * k] * xw[2 * (p0 + j)] + s.gvec[2 * k + 1] * xw[2 * (p0 + j) + 1], &sv, &cv);
    A[e] = make_double2(cv, sv);
  }
  __syncthreads();
  for (int c = 0; c < m; c++) {
    if (tid < 64) {
      // pivot row: the largest |re| + |im| among rows c..m-1 of column c
      double best = -1.0;
      int bi = c;
      for (int r = c + tid; r < m; r += 64) {
        const double2 a = A[r * m + c];
        const double v = fabs(a.x) + fabs(a.y);
        if (v > best) {
          best = v;
          bi = r;
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (tid == 0) {
        sh_piv = bi;
        perm[c] = bi;
      }
    }
    __syncthreads();
    const int pr = sh_piv;
    if (pr != c)
      for (int j = tid; j < m; j += 256) {
        const double2 t = A[c * m + j];
        A[c * m + j] = A[pr * m + j];
        A[pr * m + j] = t;
      }
    __syncthreads();
    const double2 pv = A[c * m + c];
    const double d = pv.x * pv.x + pv.y * pv.y;
    const double2 inv = make_double2(pv.x / d, -pv.y / d);
    __syncthreads();
    // row c scaled by 1/pivot, its pivot entry replaced by 1/pivot
    for (int j = tid; j < m; j += 256) {
      const double2 a = j == c ? make_double2(1.0, 0.0) : A[c * m + j];
      A[c * m + j] = make_double2(a.x * inv.x - a.y * inv.y, a.x * inv.y + a.y * inv.x);
    }
    __syncthreads();
    // column c before the rows change it
    if (tid < m) colc[tid] = A[tid * m + c];
    __syncthreads();
    // the other rows: row r -= A(r, c) row c, with A(r, c) -> -A(r, c) / pivot
    for (int e = tid; e < m * m; e += 256) {
      const int r = e / m, j = e % m;
      if (r == c) continue;
      const double2 f = colc[r];
      const double2 u = A[c * m + j];
      if (j == c) A[e] = make_double2(-(f.x * u.x - f.y * u.y), -(f.x * u.y + f.y * u.x));
      else A[e] = make_double2(A[e].x - (f.x * u.x - f.y * u.y), A[e].y - (f.x * u.y + f.y * u.x));
    }
    __syncthreads();
  }
  // the row swaps of A are column swaps of the inverse, undone last first
  for (int c = m - 1; c >= 0; c--) {
    const int pr = perm[c];
    if (pr != c)
      for (int r = tid; r < m; r += 256) {
        const double2 t = A[r * m + c];
        A[r * m + c] = A[r * m + pr];
        A[r * m + pr] = t;
      }
    __syncthreads();
  }
  typedef typename SemZ<Z>::R R;
  Z* out = ainv + ((int64_t)w * 2 + sp) * mm * mm;
  for (int e = tid; e < m * m; e += 256) {
    const int r = e / m, j = e % m;
    out[r * mm + j] = Z{(R)A[e].x, (R)A[e].y};
  }
}

// Moves of electrons j0..j1-1, in order: one 256-thread workgroup per
// walker, holding in LDS the inverses of the spins the range touches (sp0
// and up) and the positions for the call.  Per move: every thread draws the
// same three uniforms from the walker's stream (so no broadcast is needed);
// threads 0..m-1 form the new orbital column a'; four threads per row l
// form t_l = (A^-1 a')_l while the same threads sum the Jastrow change over
// the partners; after a barrier every thread holds R = t_j' and dU and
// takes the same decision, and an accepted move updates the m x m inverse,
// 256 threads over its elements.  Two barriers per move, reached by every
// thread.
template <typename Z>
__global__ __launch_bounds__(256) void k_sem_moves(DevSys s, int j0, int j1, int sp0, double sigma,
                                                   double* __restrict__ x, int64_t* __restrict__ rng,
                                                   Z* __restrict__ ainv, int mm, double* __restrict__ dlogpsi,
                                                   double* __restrict__ nacc) {
  typedef typename SemZ<Z>::R R;
  extern __shared__ __attribute__((aligned(16))) unsigned char sem_lds[];
  const int w = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int N = s.nptot;
  const int nsp = ptype(s, j1 - 1) - sp0 + 1;
  Z* B = reinterpret_cast<Z*>(sem_lds);  // [nsp][mm][mm]
  Z* ap = B + nsp * mm * mm;             // new orbital column a'
  Z* rowj = ap + 64;                      // row j' of the inverse before the update
  Z* tv = rowj + 64;                      // t = A^-1 a'
  double* xs = reinterpret_cast<double*>(tv + 64);  // positions [N][2]
  double* red = xs + 2 * N;                         // per-wavefront Jastrow sums
  const Z* gin = ainv + ((int64_t)w * 2 + sp0) * mm * mm;
  for (int e = tid; e < nsp * mm * mm; e += 256) B[e] = gin[e];
  for (int e = tid; e < 2 * N; e += 256) xs[e] = x[(int64_t)w * 2 * N + e];
  uint64_t st = (uint64_t)rng[w];
  double dl = 0.0, na = 0.0;
  __syncthreads();
  for (int j = j0; j < j1; j++) {
    const int sp = ptype(s, j);
    const int m = pick2(s.np, sp), jj = j - pick2(s.ipf, sp);
    Z* A = B + (sp - sp0) * mm * mm;
    // the proposal: a Gaussian step of width sigma per dimension (Box-Muller)
    const double u1 = rannyu(st), u2 = rannyu(st), ua = rannyu(st);
    double sv, cv;
    qsincos(6.283185307179586 * u2, &sv, &cv);
    const double rad = sigma * sqrt(-2.0 * log(u1));
    const double xo = xs[2 * j], yo = xs[2 * j + 1];
    const double xn = xo + rad * cv, yn = yo + rad * sv;
    if (tid < m) {
      double so, co;
      qsincos(s.gvec[2 * tid] * xn + s.gvec[2 * tid + 1] * yn, &so, &co);
      ap[tid] = Z{(R)co, (R)so};
    }
    __syncthreads();
    // t_l over four threads per row
    const int l = tid >> 2, q = tid & 3;
    R tr = 0, ti = 0;
    if (l < m)
      for (int k = q; k < m; k += 4) {
        const Z a = A[l * mm + k], b = ap[k];
        tr += a.x * b.x - a.y * b.y;
        ti += a.x * b.y + a.y * b.x;
      }
    tr += __shfl_xor(tr, 1, 64);
    ti += __shfl_xor(ti, 1, 64);
    tr += __shfl_xor(tr, 2, 64);
    ti += __shfl_xor(ti, 2, 64);
    if (q == 0 && l < m) tv[l] = Z{tr, ti};
    if (tid < m) rowj[tid] = A[jj * mm + tid];
    // the Jastrow change over the partners of j
    double du = 0.0;
    for (int p = tid; p < N; p += 256) {
      if (p == j) continue;
      const TabRef tu2 = tabref<TK_U2>(s, tpair(sp, ptype(s, p)));
      if (tu2.row0 < 0) continue;
      const double xp = xs[2 * p], yp = xs[2 * p + 1];
      const PairGeo pn = pair_geo(s, xn, yn, xp, yp), po = pair_geo(s, xo, yo, xp, yp);
      double fn, fo, d1, d2;
      spline3(s, tu2, pn.ind, pn.rem, fn, d1, d2);
      spline3(s, tu2, po.ind, po.rem, fo, d1, d2);
      du += fn - fo;
    }
    du = wave_sum(du);
    if (lane == 0) red[wv] = du;
    __syncthreads();
    du = red[0] + red[1] + red[2] + red[3];
    const Z rz = tv[jj];
    const double mod2 = (double)rz.x * (double)rz.x + (double)rz.y * (double)rz.y;
    if (ua < mod2 * exp(-2.0 * du)) {
      // A^-1 -= (t - e_j') (row j' of A^-1) / R, 1/R as conj(R) / |R|^2
      const R inv2 = (R)(1.0 / mod2);
      for (int e = tid; e < m * m; e += 256) {
        const int r = e / m, k = e - r * m;
        const Z t = tv[r], b = rowj[k];
        const R vr = r == jj ? t.x - (R)1 : t.x, vi = t.y;
        const R fr = -(vr * rz.x + vi * rz.y) * inv2, fi = -(vi * rz.x - vr * rz.y) * inv2;
        Z a = A[r * mm + k];
        a.x += fr * b.x - fi * b.y;
        a.y += fr * b.y + fi * b.x;
        A[r * mm + k] = a;
      }
      if (tid == 0) {
        xs[2 * j] = xn;
        xs[2 * j + 1] = yn;
      }
      dl += 0.5 * log(mod2) - du;
      na += 1.0;
    }
  }
  __syncthreads();
  Z* gout = ainv + ((int64_t)w * 2 + sp0) * mm * mm;
  for (int e = tid; e < nsp * mm * mm; e += 256) gout[e] = B[e];
  for (int e = tid; e < 2 * N; e += 256) x[(int64_t)w * 2 * N + e] = xs[e];
  if (tid == 0) {
    rng[w] = (int64_t)st;
    dlogpsi[w] += dl;
    nacc[w] += na;
  }
}

static int sem_check(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int precision) {
  if (!c || !w || !d || w->nw < 0) return arg_fail("null argument");
  if (w->nw > 0 && (!w->x || !w->rng || !d->ainv || !d->dlogpsi || !d->nacc))
    return arg_fail("null buffer for single-electron moves");
  if (precision != QMC_PREC_FP64 && precision != QMC_PREC_MIXED) return arg_fail("unknown precision mode");
  if (!c->b_zero) return arg_fail("single-electron moves need a trial function without backflow");
  if (c->s.has_u3) return arg_fail("single-electron moves support the two-body Jastrow only");
  if (!c->s.has_det) return arg_fail("single-electron moves need the Slater determinants");
  if (std::max(c->s.np[0], c->s.ntypes > 1 ? c->s.np[1] : 0) > QMC_MAX_ORB)
    return arg_fail("single-electron moves support at most 64 particles per spin", QMC_ERR_LIMIT);
  return QMC_OK;
}

static int sem_mm(const qmc_ctx* c) { return std::max(c->s.np[0], c->s.ntypes > 1 ? c->s.np[1] : 0); }

extern "C" int qmc_sem_refresh(qmc_ctx* c, const qmc_walkers* w, const qmc_sem_state* d, int precision,
                               void* stream) {
  int rc = sem_check(c, w, d, precision);
  if (rc) return rc;
  if (w->nw == 0) return QMC_OK;
  HIPCHK(hipSetDevice(c->device));
  const int mm = sem_mm(c);
  const size_t lds = sizeof(double2) * (size_t)mm * mm;
  if (lds > 65536)
    HIPCHK(hipFuncSetAttribute(precision == QMC_PREC_MIXED ? (const void*)k_sem_inverse<float2>
                                                           : (const void*)k_sem_inverse<double2>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const dim3 grid(w->nw, c->s.ntypes);
  if (precision == QMC_PREC_MIXED)
    hipLaunchKernelGGL(k_sem_inverse<float2>, grid, dim3(256), lds, (hipStream_t)stream, c->s, w->nw, w->x,
                       (float2*)d->ainv, mm);
  else
    hipLaunchKernelGGL(k_sem_inverse<double2>, grid, dim3(256), lds, (hipStream_t)stream, c->s, w->nw, w->x,
                       (double2*)d->ainv, mm);
  HIPCHK(hipGetLastError());
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
  const int mm = sem_mm(c), N = c->s.nptot;
  const int sp0 = (c->s.ntypes > 1 && j0 >= c->s.ipf[1]) ? 1 : 0;
  const int sp1 = (c->s.ntypes > 1 && j1 - 1 >= c->s.ipf[1]) ? 1 : 0;
  const size_t zs = precision == QMC_PREC_MIXED ? sizeof(float2) : sizeof(double2);
  const size_t lds = (size_t)(sp1 - sp0 + 1) * mm * mm * zs + 3 * 64 * zs + (2 * (size_t)N + 4) * sizeof(double);
  int lds_max = 0;
  HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
  if (lds > (size_t)lds_max) return arg_fail("a walker's inverses do not fit in LDS: move one spin per call", QMC_ERR_LIMIT);
  const void* kf = precision == QMC_PREC_MIXED ? (const void*)k_sem_moves<float2> : (const void*)k_sem_moves<double2>;
  if (lds > 65536) HIPCHK(hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (precision == QMC_PREC_MIXED)
    hipLaunchKernelGGL(k_sem_moves<float2>, dim3(w->nw), dim3(256), lds, (hipStream_t)stream, c->s, j0, j1, sp0,
                       sigma, w->x, w->rng, (float2*)d->ainv, mm, d->dlogpsi, d->nacc);
  else
    hipLaunchKernelGGL(k_sem_moves<double2>, dim3(w->nw), dim3(256), lds, (hipStream_t)stream, c->s, j0, j1, sp0,
                       sigma, w->x, w->rng, (double2*)d->ainv, mm, d->dlogpsi, d->nacc);
  HIPCHK(hipGetLastError());
  return QMC_OK;
}
