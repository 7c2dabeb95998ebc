// qmc_dmc.hip -- diffusion Monte Carlo with fixed-population branching on
// the device: the reference's dmc driver (reptation.f:872-921) with branch2
// (reptation.f:2162-2271), each walker playing one MPI rank that holds one
// configuration.
//
// A step is mmstep moves of every walker (the VMC step of qmc_vmc.hip: the
// rank's own random stream), each followed by the walker's weight factor and
// the weighted averages(2), then branch2: the weights of all ranks, gathered
// in rank order, are normalized to the population and combed with the second
// generator, which every rank draws alike; each rank left with no copy then
// receives one from the lowest rank with copies to spare (the send/receive
// loop of branch2).  As in the reference's send/receive, the receiving rank
// takes the configuration, drift, Laplacian and properties, and keeps its own
// walker age and fixed-phase term (reptation.f:2279-2317 send x, g, h and
// the properties; the age and sign live in separate arrays).
#include "qmc_internal.h"

#define DMC_WMULT 10.0  // branch2's bound of a weight (reptation.f:2167-2168)

// Start of a step: every weight 1 (dmc, reptation.f:886-890).
__global__ void k_dmc_begin(int nw, double* __restrict__ wate) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw) wate[w] = 1.0;
}

// Before a move: each walker's energy and age.
__global__ void k_dmc_pre(int nw, int np_, int jetot, const double* __restrict__ props,
                          const double* __restrict__ age, double* __restrict__ eprev,
                          double* __restrict__ aprev) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  eprev[w] = props[(int64_t)w * np_ + jetot];
  aprev[w] = age[w];
}

// After move + metropolis_test: the local energy metropolis_test leaves
// (reptation.f:2630-2632, 2652-2653: the mean of the new and old energies
// when accepted, the old one otherwise), the weight factor
// peso = exp(-npnorm (elocal - etrial) delta) bounded at mmult
// (reptation.f:897-903), wate *= peso, and averages(2) with weight wate
// (reptation.f:4300-4305: blk += p_old wate, blk_norm += wate).  One
// wavefront per walker, lanes over properties.
__global__ __launch_bounds__(QMC_BLOCK) void k_dmc_weight(
    int nw, int np_, int jetot, int npnorm, double delta, double etrial, double mage, double mmult,
    const double* __restrict__ props, const double* __restrict__ age,
    const double* __restrict__ eprev, const double* __restrict__ aprev, double* __restrict__ wate,
    double* __restrict__ wsum, double* __restrict__ blk, int32_t* __restrict__ counts) {
  const int w = blockIdx.x * (QMC_BLOCK / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nw) return;
  const double* p = props + (int64_t)w * np_;
  const double eold = eprev[w];
  const double elocal = age[w] == 0.0 ? 0.5 * (p[jetot] + eold) : eold;
  double peso = exp(-(double)npnorm * (elocal - etrial) * delta);
  const bool bounded = peso > mmult;
  if (bounded) peso = mmult;
  const double wt = wate[w] * peso;
  double* b = blk + (int64_t)w * np_;
  for (int k = lane; k < np_; k += 64) b[k] += p[k] * wt;
  if (lane == 0) {
    wate[w] = wt;
    wsum[w] += wt;
    if (bounded) atomicAdd(&counts[0], 1);
    // move counts a forced move when the age reached mage (reptation.f:4650-4653)
    if (w == 0 && aprev[0] == mage) atomicAdd(&counts[3], 1);
  }
}

// branch2 (reptation.f:2162-2271) over the nw ranks.  The sums of the
// weights and of their fractional parts, which the reference forms in rank
// order over the gathered values, and the comb run on one thread over LDS
// copies, in the reference's order of operations; the bound, the
// normalization, the integer parts and the transfers (integer prefix sums)
// are spread over the workgroup.  LDS: nw doubles (the weights, then their
// fractional parts, then the prefix sums of the extra copies) and nw ints
// (the multiplicities).  On return src[j] is the walker j takes its
// configuration from.
#define DMC_BRANCH_THREADS 1024

__global__ __launch_bounds__(DMC_BRANCH_THREADS) void k_dmc_branch(int nw, double* __restrict__ wate,
                                                                   int64_t* __restrict__ rng2,
                                                                   int32_t* __restrict__ src,
                                                                   int32_t* __restrict__ counts) {
  extern __shared__ __attribute__((aligned(16))) unsigned char dmc_lds[];
  double* sw = reinterpret_cast<double*>(dmc_lds);
  int* sm = reinterpret_cast<int*>(sw + nw);
  __shared__ double sh_somma;
  __shared__ int pe[DMC_BRANCH_THREADS], pr[DMC_BRANCH_THREADS];
  const int tid = threadIdx.x;
  // weights above wmult are bounded (reptation.f:2176-2182)
  for (int i = tid; i < nw; i += DMC_BRANCH_THREADS) {
    double v = wate[i];
    if (v > DMC_WMULT) {
      v = DMC_WMULT;
      atomicAdd(&counts[1], 1);
      if (i == 0) atomicAdd(&counts[2], 1);
    }
    sw[i] = v;
  }
  __syncthreads();
  if (tid == 0) {
    double somma = 0.0;
#pragma unroll 8
    for (int i = 0; i < nw; i++) somma = somma + sw[i];
    sh_somma = somma;
  }
  __syncthreads();
  // multiplicities: integer parts, and the fractional parts to comb
  // (reptation.f:2189-2195)
  const double somma = sh_somma;
  for (int i = tid; i < nw; i += DMC_BRANCH_THREADS) {
    const double ww = ((double)nw * sw[i]) / somma;
    const int tm = (int)ww;
    sm[i] = tm;
    sw[i] = ww - (double)tm;
  }
  __syncthreads();
  if (tid == 0) {
    double k0 = 0.0;
#pragma unroll 8
    for (int i = 0; i < nw; i++) k0 = k0 + sw[i];
    // The comb: nint(k0) draws of rannyu2, one in each unit of the running
    // sum of the fractional parts (reptation.f:2196-2210).  The reference
    // moves its walker pointer while k < r (then marks the walker), then
    // while k < 1, and subtracts 1 from k before each draw; walked here in
    // walker order, k gets the same additions and subtractions in the same
    // order: after each walker's fractional part, the draws whose points
    // that walker reaches are taken (marked, k reaches 1, k - 1 and the next
    // draw), so the loads do not wait on the comparisons.
    const long ndraw = lround(k0);
    uint64_t st = (uint64_t)rng2[0];
    long d = 0;
    double k = 0.0, r = 0.0;
    bool seek_r = false;
    if (ndraw > 0) {
      r = rannyu(st);
      d = 1;
      seek_r = true;  // k = 1 - 1 before the first draw
    }
    // eight fractional parts at a time into registers, the marks counted
    // in registers and added to the multiplicities per chunk
    for (int j0 = 0; j0 < nw && d > 0; j0 += 8) {
      double v[8];
      int mk[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        v[u] = j0 + u < nw ? sw[j0 + u] : 0.0;
        mk[u] = 0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (j0 + u < nw && d > 0) {
          k = k + v[u];
          for (;;) {
            if (seek_r) {
              if (k < r) break;
              mk[u] += 1;
              seek_r = false;
            }
            if (k < 1.0) break;
            if (d >= ndraw) {
              d = 0;
              break;
            }
            k = k - 1.0;
            r = rannyu(st);
            d++;
            seek_r = true;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (mk[u]) sm[j0 + u] += mk[u];
    }
    rng2[0] = (int64_t)st;
  }
  __syncthreads();
  // The ranks left with no copy take the extra copies of the others: each
  // transfer from the lowest rank with more than one to the lowest rank with
  // none (reptation.f:2220-2258).  So receiver q (in walker order) takes
  // from the walker s whose extra copies, counted in walker order, cover q:
  // E(s) <= q < E(s+1), E the exclusive prefix sum of max(mult - 1, 0).
  const int chunk = (nw + DMC_BRANCH_THREADS - 1) / DMC_BRANCH_THREADS;
  const int i0 = min(nw, tid * chunk), i1 = min(nw, i0 + chunk);
  int ne = 0, nr = 0;
  for (int i = i0; i < i1; i++) {
    ne += max(sm[i] - 1, 0);
    nr += sm[i] == 0 ? 1 : 0;
  }
  pe[tid] = ne;
  pr[tid] = nr;
  __syncthreads();
  for (int o = 1; o < DMC_BRANCH_THREADS; o <<= 1) {
    const int ae = tid >= o ? pe[tid - o] : 0, ar = tid >= o ? pr[tid - o] : 0;
    __syncthreads();
    pe[tid] += ae;
    pr[tid] += ar;
    __syncthreads();
  }
  // exclusive prefix sums of the extras per walker, in the fractional parts'
  // place (no longer read); E(nw) is the total
  double* E = sw;
  int e = pe[tid] - ne, q = pr[tid] - nr;
  for (int i = i0; i < i1; i++) {
    E[i] = (double)e;
    e += max(sm[i] - 1, 0);
  }
  const int etot = pe[DMC_BRANCH_THREADS - 1];
  __syncthreads();
  for (int i = i0; i < i1; i++) {
    if (sm[i] >= 1) {
      src[i] = i;
      continue;
    }
    int sidx = i;
    if (q < etot) {
      // the last walker with E <= q
      int lo = 0, hi = nw - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)E[mid] <= q) lo = mid;
        else hi = mid - 1;
      }
      sidx = lo;
    }
    src[i] = sidx;
    q++;
  }
}

static size_t dmc_branch_lds(int nw) { return (size_t)nw * (sizeof(double) + sizeof(int)); }

// The copies the branching asks for: walker j takes the configuration, drift,
// Laplacian and properties of walker src[j].  The walkers copied from keep
// their own (a multiplicity of at least 1), so no walker is both read and
// written.  One wavefront per walker.
__global__ __launch_bounds__(QMC_BLOCK) void k_dmc_copy(int nw, int N, int np_, const int32_t* __restrict__ src,
                                                        double* __restrict__ x, double* __restrict__ g,
                                                        double* __restrict__ h, double* __restrict__ props) {
  const int w = blockIdx.x * (QMC_BLOCK / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nw) return;
  const int from = src[w];
  if (from == w || from < 0 || from >= nw) return;
  for (int k = lane; k < 2 * N; k += 64) {
    x[(int64_t)w * 2 * N + k] = x[(int64_t)from * 2 * N + k];
    g[(int64_t)w * 2 * N + k] = g[(int64_t)from * 2 * N + k];
  }
  if (lane < 2) h[(int64_t)w * 2 + lane] = h[(int64_t)from * 2 + lane];
  for (int k = lane; k < np_; k += 64) props[(int64_t)w * np_ + k] = props[(int64_t)from * np_ + k];
}

static int check_dmc(qmc_ctx* c, int nw, const qmc_dmc_state* d) {
  if (!c || !d || nw < 0) return arg_fail("bad DMC state");
  if (nw > 0 && (!d->wate || !d->wsum || !d->src || !d->rng2 || !d->counts))
    return arg_fail("null DMC buffer");
  return QMC_OK;
}

// The branching's LDS holds 12 bytes per walker: the populations the LDS of
// a workgroup can take (13000 walkers at 160 KB).
static int dmc_branch_check(qmc_ctx* c, int nw) {
  int lds_max = 0;
  HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
  const size_t need = dmc_branch_lds(nw) + 64;
  if (need > (size_t)lds_max) return arg_fail("DMC population above what the branching's LDS holds", QMC_ERR_LIMIT);
  if (need > 65536)
    HIPCHK(hipFuncSetAttribute((const void*)k_dmc_branch, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)dmc_branch_lds(nw)));
  return QMC_OK;
}

extern "C" int qmc_dmc_branch(qmc_ctx* c, int nw, double* d_wate, int64_t* d_rng2, int32_t* d_src,
                              int32_t* d_counts, void* stream) {
  if (!c || nw < 0 || (nw > 0 && (!d_wate || !d_rng2 || !d_src || !d_counts)))
    return arg_fail("bad branching arguments");
  if (nw == 0) return QMC_OK;
  HIPCHK(hipSetDevice(c->device));
  int rc;
  rc = dmc_branch_check(c, nw);
  if (rc) return rc;
  hipLaunchKernelGGL(k_dmc_branch, dim3(1), dim3(DMC_BRANCH_THREADS), dmc_branch_lds(nw), (hipStream_t)stream,
                     nw, d_wate, d_rng2, d_src, d_counts);
  HIPCHK(hipGetLastError());
  return QMC_OK;
}

extern "C" int qmc_dmc_sweep(qmc_ctx* c, const qmc_walkers* w, const qmc_dmc_state* d, int nsteps,
                             const qmc_dmc_params* prm, void* stream) {
  if (!c || !w || !prm) return arg_fail("null argument");
  int rc = check_dmc(c, w->nw, d);
  if (rc) return rc;
  if (w->nw > 0 && (!w->x || !w->g || !w->h || !w->props || !w->grad2_ph || !w->age || !w->rng))
    return arg_fail("null walker buffer");
  if (w->nw > 0 && !w->blk) return arg_fail("DMC needs the block sums (w->blk)");
  if (w->nw == 0 || nsteps <= 0) return QMC_OK;
  if (!(prm->delta > 0.0)) return arg_fail("delta must be positive");
  if (prm->mmstep < 1) return arg_fail("mmstep must be at least 1");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  rc = dmc_branch_check(c, w->nw);
  if (rc) return rc;
  rc = ensure_vmc_scratch(c, w->nw);
  if (rc) return rc;
  rc = ensure_ws(c, w->nw);
  if (rc) return rc;
  const DevSys& s = c->s;
  c->delta = prm->delta;
  // sigma (reptation.f:855-870)
  const double var0 = 2.0 * s.hbs2m[0] * prm->delta;
  const double var1 = s.ntypes > 1 ? 2.0 * s.hbs2m[1] * prm->delta : 1.0;
  const int nw = w->nw, NP = c->info.n_props, jetot = c->info.jetot - 1;
  // the moves add the weighted block sums here, not the plain ones
  qmc_walkers wv = *w;
  wv.blk = nullptr;
  const int wpb = QMC_BLOCK / 64;
  for (int k = 0; k < nsteps; k++) {
    hipLaunchKernelGGL(k_dmc_begin, dim3((nw + 255) / 256), dim3(256), 0, st, nw, d->wate);
    for (int m = 0; m < prm->mmstep; m++) {
      hipLaunchKernelGGL(k_dmc_pre, dim3((nw + 255) / 256), dim3(256), 0, st, nw, NP, jetot, w->props, w->age,
                         c->d_eprev, c->d_aprev);
      rc = vmc_step(c, &wv, st, prm->delta, var0, var1, prm->mage, prm->bm_order, nullptr, nullptr);
      if (rc) return rc;
      hipLaunchKernelGGL(k_dmc_weight, dim3((nw + wpb - 1) / wpb), dim3(QMC_BLOCK), 0, st, nw, NP, jetot,
                         s.npnorm, prm->delta, prm->etrial, prm->mage, (double)prm->mmult, w->props, w->age,
                         c->d_eprev, c->d_aprev, d->wate, d->wsum, w->blk, d->counts);
    }
    hipLaunchKernelGGL(k_dmc_branch, dim3(1), dim3(DMC_BRANCH_THREADS), dmc_branch_lds(nw), st, nw, d->wate,
                       d->rng2, d->src, d->counts);
    hipLaunchKernelGGL(k_dmc_copy, dim3((nw + wpb - 1) / wpb), dim3(QMC_BLOCK), 0, st, nw, s.nptot, NP, d->src,
                       w->x, w->g, w->h, w->props);
  }
  HIPCHK(hipGetLastError());
  return QMC_OK;
}
