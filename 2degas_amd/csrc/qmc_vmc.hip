// qmc_vmc.hip -- batched VMC step on the device: drift-diffusion proposal,
// evaluation, Metropolis test and block accumulation for every walker.
//
// Each walker owns a stream of the reference's 48-bit LCG (rannyu,
// reptation.f:5247-5291), so a batch of nw walkers draws exactly what nw MPI
// ranks of the reference draw (rank r is seeded with seed(4) = 2*(seed(4)+r)+1,
// reptation.f:37-39).
#include "qmc_internal.h"

#include <cstring>

// move, proposal half (reptation.f:4630-4646): for every moving type,
// particle and dimension chi = cos(pi u1) sqrt(-2 var log u2),
// x_new = chi - var g_old + x_old, w = -sum chi^2 / var.
// One wavefront per walker, lanes over coordinates.  The reference draws two
// uniforms per moved coordinate in coordinate order, so the coordinate with
// ordinal o among the moved ones takes draws 2o+1 and 2o+2 of the walker's
// stream: each lane jumps there with a^(2o+1).  The chi^2 terms go to LDS
// and lane 0 adds them in the reference's order.
__global__ __launch_bounds__(QMC_BLOCK) void k_vmc_propose(
    DevSys s, int nw, const double* __restrict__ x_old, const double* __restrict__ g_old,
    double* __restrict__ x_new, int64_t* __restrict__ rng, double* __restrict__ wold,
    double var0, double var1, int bm_order) {
  __shared__ double chi2[QMC_BLOCK / 64][2 * QMC_MAX_PARTICLES];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int w = blockIdx.x * (QMC_BLOCK / 64) + wv;
  if (w >= nw) return;  // no barrier below: each wavefront works alone
  const int N = s.nptot, n2 = 2 * N;
  const double* xo = x_old + (int64_t)w * n2;
  const double* go = g_old + (int64_t)w * n2;
  double* xn = x_new + (int64_t)w * n2;
  const uint64_t st0 = (uint64_t)rng[w];
  const double pi = 3.141592653589793;
  const uint64_t msk = (1ULL << 48) - 1ULL, a = 34522712143931ULL;
  const bool two = s.ntypes > 1;
  const bool mov0 = s.hbs2m[0] != 0.0, mov1 = two && s.hbs2m[1] != 0.0;
  const int m0 = mov0 ? 2 * s.np[0] : 0;  // moved coordinates of type 0
  const int nmov = m0 + (mov1 ? 2 * s.np[1] : 0);
  double* c2 = chi2[wv];
  for (int e = lane; e < n2; e += 64) {
    const int t = two && e >= 2 * s.ipf[1] ? 1 : 0;
    if (!(t ? mov1 : mov0)) {
      xn[e] = xo[e];
      continue;
    }
    const int o = t ? m0 + e - 2 * s.ipf[1] : e - 2 * s.ipf[0];
    const double var = t ? var1 : var0;
    const uint64_t x1 = (lcg_pow(2 * (uint64_t)o + 1) * st0) & msk, x2 = (a * x1) & msk;
    const double u1 = (double)x1 * 0x1p-48, u2 = (double)x2 * 0x1p-48;
    const double ua = bm_order == 0 ? u1 : u2, ub = bm_order == 0 ? u2 : u1;
    const double chi = cos(pi * ua) * sqrt(-2.0 * var * log(ub));
    c2[e] = chi * chi;
    xn[e] = chi + (-var * go[e]) + xo[e];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane == 0) {
    double wo = 0.0;
    for (int t = 0; t < s.ntypes; t++) {
      if (!(t ? mov1 : mov0)) continue;
      const double var = t == 0 ? var0 : var1;
      const int e0 = 2 * pick2(s.ipf, t), e1 = e0 + 2 * pick2(s.np, t);
      double wt = 0.0;
      for (int e = e0; e < e1; e++) wt = wt - c2[e];
      wo = wo + wt * (1.0 / var);
    }
    wold[w] = wo;
    rng[w] = (int64_t)((lcg_pow(2 * (uint64_t)nmov) * st0) & msk);
  }
}

// move, acceptance half (reptation.f:4648-4678) + metropolis_test
// (reptation.f:2622-2659) + averages(2) (reptation.f:4300-4305).
// One wavefront per walker.
__global__ __launch_bounds__(QMC_BLOCK) void k_vmc_accept(
    DevSys s, int nw, double var0, double var1, double mage, double* __restrict__ x_old,
    double* __restrict__ g_old, double* __restrict__ h_old, double* __restrict__ p_old,
    double* __restrict__ age, const double* __restrict__ x_new, const double* __restrict__ g_new,
    const double* __restrict__ h_new, const double* __restrict__ p_new, int64_t* __restrict__ rng,
    const double* __restrict__ wold, const int32_t* __restrict__ status,
    int32_t* __restrict__ sticky, double* __restrict__ blk, double* __restrict__ trace_p,
    double* __restrict__ trace_age, int jacc, int jltf) {
  const int w = blockIdx.x * (QMC_BLOCK / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nw) return;
  const int N = s.nptot, NP = s.n_props;
  const double* xo = x_old + (int64_t)w * 2 * N;
  const double* xn = x_new + (int64_t)w * 2 * N;
  const double* gn = g_new + (int64_t)w * 2 * N;
  double wt[2] = {0.0, 0.0};
  for (int ip = lane; ip < N; ip += 64) {
    const int t = ptype(s, ip);
    const double var = t == 0 ? var0 : var1;
    for (int d = 0; d < 2; d++) {
      const double drift = -var * gn[2 * ip + d];
      const double chi = xo[2 * ip + d] - xn[2 * ip + d] - drift;
      if (t == 0) wt[0] = wt[0] - chi * chi;
      else wt[1] = wt[1] - chi * chi;
    }
  }
  wt[0] = wave_sum(wt[0]);
  wt[1] = wave_sum(wt[1]);
  double p = 0.0;
  int acc = 0;
  if (lane == 0) {
    const double* po = p_old + (int64_t)w * NP;
    const double* pn = p_new + (int64_t)w * NP;
    if (age[w] == mage) {
      p = 1.0;
    } else {
      double wnew = 0.0;
      if (s.hbs2m[0] != 0.0) wnew = wnew + wt[0] * (1.0 / var0);
      if (s.ntypes > 1 && s.hbs2m[1] != 0.0) wnew = wnew + wt[1] * (1.0 / var1);
      p = (po[jltf] - pn[jltf]) * 2.0 + 0.5 * (wnew - wold[w]);
      p = fmin(0.0, fmax(-50.0, p));
      p = exp(p);
    }
    double u = 1.0;
    if (p < 1.0) {
      uint64_t st = (uint64_t)rng[w];
      u = rannyu(st);
      rng[w] = (int64_t)st;
    }
    acc = p >= u;
    sticky[w] |= status[w];
  }
  acc = __shfl(acc, 0, 64);
  p = __shfl(p, 0, 64);
  // copy new into old where accepted, set p_old(jacc) = p and add p_old to
  // the block sums; each lane owns its slots, so no lane reads another's
  double* po = p_old + (int64_t)w * NP;
  const double* pn = p_new + (int64_t)w * NP;
  double* b = blk ? blk + (int64_t)w * NP : nullptr;
  for (int k = lane; k < NP; k += 64) {
    const double v = (k == jacc) ? p : (acc ? pn[k] : po[k]);
    po[k] = v;
    if (b) b[k] += v;
  }
  if (acc) {
    double* xo2 = x_old + (int64_t)w * 2 * N;
    double* go = g_old + (int64_t)w * 2 * N;
    for (int k = lane; k < 2 * N; k += 64) {
      xo2[k] = xn[k];
      go[k] = gn[k];
    }
    if (lane < 2) h_old[(int64_t)w * 2 + lane] = h_new[(int64_t)w * 2 + lane];
  }
  if (lane == 0) {
    const double a = acc ? 0.0 : age[w] + 1.0;
    age[w] = a;
    if (trace_p) trace_p[w] = p;
    if (trace_age) trace_age[w] = a;
  }
}

// read_conf's start: age 0 and p_old(jacc) = 1 after the evaluation.
__global__ void k_vmc_start(int nw, int np_, int jacc, double* __restrict__ props,
                            double* __restrict__ age, int32_t* __restrict__ sticky) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  props[(int64_t)w * np_ + jacc] = 1.0;
  age[w] = 0.0;
  sticky[w] = 0;
}

void free_vmc(qmc_ctx* c) {
  free_vmc_graph(c);
  (void)hipFree(c->d_xn);
  (void)hipFree(c->d_gn);
  (void)hipFree(c->d_hn);
  (void)hipFree(c->d_pn);
  (void)hipFree(c->d_wold);
  (void)hipFree(c->d_stn);
  (void)hipFree(c->d_sticky);
  (void)hipFree(c->d_eprev);
  (void)hipFree(c->d_aprev);
  c->d_xn = c->d_gn = c->d_hn = c->d_pn = c->d_wold = nullptr;
  c->d_stn = c->d_sticky = nullptr;
  c->d_eprev = c->d_aprev = nullptr;
  c->vmc_nw = 0;
}

// Proposal and status buffers for nw walkers (or reptiles).
int ensure_vmc_scratch(qmc_ctx* c, int nw) {
  if (nw <= c->vmc_nw) return QMC_OK;
  free_vmc(c);
  const size_t N = c->s.nptot, NP = c->info.n_props;
  HIPCHK(hipMalloc(&c->d_xn, sizeof(double) * 2 * N * nw));
  HIPCHK(hipMalloc(&c->d_gn, sizeof(double) * 2 * N * nw));
  HIPCHK(hipMalloc(&c->d_hn, sizeof(double) * 2 * nw));
  HIPCHK(hipMalloc(&c->d_pn, sizeof(double) * NP * nw));
  HIPCHK(hipMalloc(&c->d_wold, sizeof(double) * nw));
  HIPCHK(hipMalloc(&c->d_stn, sizeof(int32_t) * nw));
  HIPCHK(hipMalloc(&c->d_sticky, sizeof(int32_t) * nw));
  HIPCHK(hipMalloc(&c->d_eprev, sizeof(double) * nw));
  HIPCHK(hipMalloc(&c->d_aprev, sizeof(double) * nw));
  HIPCHK(hipMemset(c->d_pn, 0, sizeof(double) * NP * nw));
  // a sweep may run without a start call (state loaded by the caller)
  HIPCHK(hipMemset(c->d_sticky, 0, sizeof(int32_t) * nw));
  c->vmc_nw = nw;
  return QMC_OK;
}

static int check_walkers(qmc_ctx* c, const qmc_walkers* w) {
  if (!c || !w || w->nw < 0) return arg_fail("bad walker set");
  if (w->nw > 0 && (!w->x || !w->g || !w->h || !w->props || !w->grad2_ph || !w->age || !w->rng))
    return arg_fail("null walker buffer");
  return QMC_OK;
}

extern "C" int qmc_vmc_start(qmc_ctx* c, const qmc_walkers* w, void* stream) {
  int rc = check_walkers(c, w);
  if (rc) return rc;
  if (w->nw == 0) return QMC_OK;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  rc = ensure_vmc_scratch(c, w->nw);
  if (rc) return rc;
  rc = launch_eval(c, w->nw, w->x, w->props, w->g, w->h, w->grad2_ph, c->d_stn, st, c->delta);
  if (rc) return rc;
  hipLaunchKernelGGL(k_vmc_start, dim3((w->nw + 255) / 256), dim3(256), 0, st, w->nw,
                     c->info.n_props, c->info.jacc - 1, w->props, w->age, c->d_sticky);
  HIPCHK(hipGetLastError());
  return QMC_OK;
}

// One VMC step for every walker: proposal, evaluation, acceptance.
int vmc_step(qmc_ctx* c, const qmc_walkers* w, hipStream_t st, double delta, double var0, double var1,
             double mage, int bm_order, double* tp, double* ta) {
  const DevSys& s = c->s;
  const int nw = w->nw;
  const int wpb = QMC_BLOCK / 64;
  hipLaunchKernelGGL(k_vmc_propose, dim3((nw + wpb - 1) / wpb), dim3(QMC_BLOCK), 0, st, s, nw,
                     w->x, w->g, c->d_xn, w->rng, c->d_wold, var0, var1, bm_order);
  // the fixed-phase scratch is shared by proposals, as the COMMON is
  int rc = launch_eval(c, nw, c->d_xn, c->d_pn, c->d_gn, c->d_hn, w->grad2_ph, c->d_stn, st, delta);
  if (rc) return rc;
  hipLaunchKernelGGL(k_vmc_accept, dim3((nw + wpb - 1) / wpb), dim3(QMC_BLOCK), 0, st, s, nw,
                     var0, var1, mage, w->x, w->g, w->h, w->props, w->age, c->d_xn, c->d_gn,
                     c->d_hn, c->d_pn, w->rng, c->d_wold, c->d_stn, c->d_sticky, w->blk, tp, ta,
                     c->info.jacc - 1, c->info.jltf - 1);
  return QMC_OK;
}

void free_vmc_graph(qmc_ctx* c) {
  if (c->vmc_exec) (void)hipGraphExecDestroy(c->vmc_exec);
  c->vmc_exec = nullptr;
  c->vmc_key.clear();
  if (c->rmc_exec) (void)hipGraphExecDestroy(c->rmc_exec);
  c->rmc_exec = nullptr;
  c->rmc_key.clear();
}

extern "C" int qmc_vmc_sweep(qmc_ctx* c, const qmc_walkers* w, int nsteps, double delta,
                             double mage, int bm_order, double* trace_p, double* trace_age,
                             void* stream) {
  int rc = check_walkers(c, w);
  if (rc) return rc;
  if (w->nw == 0 || nsteps <= 0) return QMC_OK;
  if (!(delta > 0.0)) return arg_fail("delta must be positive");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  rc = ensure_vmc_scratch(c, w->nw);
  if (rc) return rc;
  rc = ensure_ws(c, w->nw);
  if (rc) return rc;
  const DevSys& s = c->s;
  // the vmc record sets the COMMON delta before sigma (reptation.f:740-746)
  c->delta = delta;
  // sigma (reptation.f:855-870)
  const double var0 = 2.0 * s.hbs2m[0] * delta;
  const double var1 = s.ntypes > 1 ? 2.0 * s.hbs2m[1] * delta : 1.0;
  const int nw = w->nw;
  const bool graph = c->use_graph && !trace_p && !trace_age && !c->timing && nsteps > 1;
  if (!graph) {
    for (int k = 0; k < nsteps; k++) {
      rc = vmc_step(c, w, st, delta, var0, var1, mage, bm_order,
                    trace_p ? trace_p + (int64_t)k * nw : nullptr,
                    trace_age ? trace_age + (int64_t)k * nw : nullptr);
      if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    return QMC_OK;
  }
  // The step as a captured graph, replayed nsteps times: rebuilt whenever
  // anything its kernels were given changes (buffers, step parameters, the
  // g-vectors).  Captured on the context's own stream, launched on st.
  double dk[4] = {var0, var1, mage, delta};
  std::vector<uint64_t> key = {(uint64_t)nw, (uint64_t)(uintptr_t)w->x, (uint64_t)(uintptr_t)w->g,
                               (uint64_t)(uintptr_t)w->h, (uint64_t)(uintptr_t)w->props,
                               (uint64_t)(uintptr_t)w->age, (uint64_t)(uintptr_t)w->rng,
                               (uint64_t)(uintptr_t)w->blk, (uint64_t)(uintptr_t)w->grad2_ph,
                               (uint64_t)(uintptr_t)c->d_xn, (uint64_t)(uintptr_t)c->d_ws,
                               (uint64_t)(uintptr_t)c->d_F, (uint64_t)(uintptr_t)c->d_A,
                               (uint64_t)(uintptr_t)s.gvec, (uint64_t)s.ng, (uint64_t)bm_order};
  for (double v : dk) {
    uint64_t b;
    memcpy(&b, &v, sizeof b);
    key.push_back(b);
  }
  if (!c->vmc_exec || key != c->vmc_key) {
    free_vmc_graph(c);
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = vmc_step(c, w, c->stream, delta, var0, var1, mage, bm_order, nullptr, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    const hipError_t e2 = hipGraphInstantiate(&c->vmc_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      c->vmc_exec = nullptr;
      return hip_fail(e2, "hipGraphInstantiate");
    }
    c->vmc_key = key;
  }
  for (int k = 0; k < nsteps; k++) HIPCHK(hipGraphLaunch(c->vmc_exec, st));
  return QMC_OK;
}

extern "C" int qmc_vmc_status(qmc_ctx* c, int nw, int32_t* sticky_host) {
  if (!c || nw < 0 || nw > c->vmc_nw || (nw > 0 && !sticky_host)) return arg_fail("bad status query");
  if (nw == 0) return QMC_OK;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(sticky_host, c->d_sticky, sizeof(int32_t) * nw, hipMemcpyDeviceToHost));
  for (int k = 0; k < nw; k++)
    if (sticky_host[k] & 2) return arg_fail("slater_bckflw_orbitals: singular orbital matrix", QMC_ERR_SINGULAR);
  return QMC_OK;
}
