// qmc_vmc.hip -- batched VMC step on the device: drift-diffusion proposal,
// evaluation, Metropolis test and block accumulation for every walker.
//
// Each walker owns a stream of the reference's 48-bit LCG (rannyu,
// reptation.f:5247-5291), so a batch of nw walkers draws exactly what nw MPI
// ranks of the reference draw (rank r is seeded with seed(4) = 2*(seed(4)+r)+1,
// reptation.f:37-39).
#include "qmc_internal.h"

// move, proposal half (reptation.f:4630-4646): for every moving type,
// particle and dimension chi = cos(pi u1) sqrt(-2 var log u2),
// x_new = chi - var g_old + x_old, w = -sum chi^2 / var.
__global__ void k_vmc_propose(DevSys s, int nw, const double* __restrict__ x_old,
                              const double* __restrict__ g_old, double* __restrict__ x_new,
                              int64_t* __restrict__ rng, double* __restrict__ wold, double var0,
                              double var1, int bm_order) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  uint64_t st = (uint64_t)rng[w];
  const double pi = 3.141592653589793;
  const int N = s.nptot;
  const double* xo = x_old + (int64_t)w * 2 * N;
  const double* go = g_old + (int64_t)w * 2 * N;
  double* xn = x_new + (int64_t)w * 2 * N;
  double wo = 0.0;
  for (int t = 0; t < s.ntypes; t++) {
    const double var = t == 0 ? var0 : var1;
    if (pick2(s.hbs2m, t) == 0.0) continue;
    double wt = 0.0;
    const int i0 = pick2(s.ipf, t), i1 = i0 + pick2(s.np, t);
    for (int ip = i0; ip < i1; ip++)
      for (int d = 0; d < 2; d++) {
        double ua, ub;
        if (bm_order == 0) { ua = rannyu(st); ub = rannyu(st); }
        else { ub = rannyu(st); ua = rannyu(st); }
        const double chi = cos(pi * ua) * sqrt(-2.0 * var * log(ub));
        wt = wt - chi * chi;
        const double drift = -var * go[2 * ip + d];
        xn[2 * ip + d] = chi + drift + xo[2 * ip + d];
      }
    wo = wo + wt * (1.0 / var);
  }
  rng[w] = (int64_t)st;
  wold[w] = wo;
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
  (void)hipFree(c->d_xn);
  (void)hipFree(c->d_gn);
  (void)hipFree(c->d_hn);
  (void)hipFree(c->d_pn);
  (void)hipFree(c->d_wold);
  (void)hipFree(c->d_stn);
  (void)hipFree(c->d_sticky);
  c->d_xn = c->d_gn = c->d_hn = c->d_pn = c->d_wold = nullptr;
  c->d_stn = c->d_sticky = nullptr;
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
  const DevSys& s = c->s;
  // the vmc record sets the COMMON delta before sigma (reptation.f:740-746)
  c->delta = delta;
  // sigma (reptation.f:855-870)
  const double var0 = 2.0 * s.hbs2m[0] * delta;
  const double var1 = s.ntypes > 1 ? 2.0 * s.hbs2m[1] * delta : 1.0;
  const int nw = w->nw;
  for (int k = 0; k < nsteps; k++) {
    hipLaunchKernelGGL(k_vmc_propose, dim3((nw + 63) / 64), dim3(64), 0, st, s, nw, w->x, w->g,
                       c->d_xn, w->rng, c->d_wold, var0, var1, bm_order);
    // the fixed-phase scratch is shared by proposals, as the COMMON is
    rc = launch_eval(c, nw, c->d_xn, c->d_pn, c->d_gn, c->d_hn, w->grad2_ph, c->d_stn, st, delta);
    if (rc) return rc;
    const int wpb = QMC_BLOCK / 64;
    hipLaunchKernelGGL(k_vmc_accept, dim3((nw + wpb - 1) / wpb), dim3(QMC_BLOCK), 0, st, s, nw,
                       var0, var1, mage, w->x, w->g, w->h, w->props, w->age, c->d_xn, c->d_gn,
                       c->d_hn, c->d_pn, w->rng, c->d_wold, c->d_stn, c->d_sticky, w->blk,
                       trace_p ? trace_p + (int64_t)k * nw : nullptr,
                       trace_age ? trace_age + (int64_t)k * nw : nullptr, c->info.jacc - 1,
                       c->info.jltf - 1);
  }
  HIPCHK(hipGetLastError());
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
