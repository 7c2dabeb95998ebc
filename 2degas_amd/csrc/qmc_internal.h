// qmc_internal.h -- host-side context shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "qmc_dev.h"
#include "qmc_hip.h"

extern thread_local std::string g_last_error;

inline int hip_fail(hipError_t e, const char* what) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  g_last_error = buf;
  return QMC_ERR_HIP;
}
#define RMC_MAX_MDELTA 64  // largest number of beads one reptation move may grow

#define HIPCHK(call)                                       \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return hip_fail(e_, #call);      \
  } while (0)

inline int arg_fail(const char* msg, int code = QMC_ERR_ARG) {
  g_last_error = msg;
  return code;
}

struct qmc_ctx {
  int device = 0;
  DevSys s{};
  qmc_info info{};
  double4* d_tab = nullptr;
  float4* d_tab32 = nullptr;  // the tables rounded to FP32 (mixed precision)
  double* d_kvec = nullptr;
  double* d_vk = nullptr;
  double* d_gvec = nullptr;
  double* d_kxyv = nullptr;  // distinct kx then ky components
  int* d_kxy = nullptr;
  TpInfo* d_tpi = nullptr;
  // workspace sized for ws_nw walkers
  int ws_nw = 0;
  // device bytes of the uploaded tables (not counting the g-vectors)
  size_t table_bytes = 0;
  double* d_ws = nullptr;
  double* d_F = nullptr;
  double* d_A = nullptr;
  WsLayout L{};
  // staging for the host-pointer entry point
  int st_nw = 0;
  double *d_x = nullptr, *d_props = nullptr, *d_g = nullptr, *d_h = nullptr, *d_g2 = nullptr;
  int32_t* d_status = nullptr;
  hipStream_t stream = nullptr;
  double delta = 0.0;  // current time step (COMMON delta) for the drift cutoff
  int kc = 16;
  size_t lds_pairs = 0, lds_det = 0, lds_cross = 0;
  size_t lds_pairs_mix = 0;
  int kchunk_mix = 1;
  int tpp_det = 0;
  bool use_gj = false;   // matrix-core Gauss-Jordan determinant path (else k_det)
  bool b_zero = true;    // every backflow table is zero (quasi-coordinates = positions)
  DevSys* d_sys = nullptr;  // device copy of s for kernels that read it from memory
  DevSys sys_sent{};        // what d_sys holds (valid when sys_valid)
  // single-electron move kernels: threads per inverse row, per precision
  // (FP64, FP32) and row class (16, 32, 64 rows)
  int sem_cg[2][3] = {{4, 2, 4}, {4, 2, 4}};
  bool sys_valid = false;
  bool gjm_part = false; // a spin below 49 particles: k_gjm_part, which skips the tiles beyond m
  size_t lds_gjm = 0;
  size_t lds_fgrad = 0;
  // per-kernel event timing
  bool timing = false;
  std::vector<hipEvent_t> ev;  // TIMING_EVENTS events per recorded launch
  int ev_used = 0;
  // VMC scratch: proposal buffers for vmc_nw walkers
  int vmc_nw = 0;
  double *d_xn = nullptr, *d_gn = nullptr, *d_hn = nullptr, *d_pn = nullptr, *d_wold = nullptr;
  int32_t *d_stn = nullptr, *d_sticky = nullptr;
  // DMC: each walker's energy and age before its latest move
  double *d_eprev = nullptr, *d_aprev = nullptr;
  // the VMC step as a captured graph, and what it was captured with
  bool use_graph = true;
  hipGraphExec_t vmc_exec = nullptr;
  std::vector<uint64_t> vmc_key;
  // the reptation move as a captured graph, and what it was captured with
  hipGraphExec_t rmc_exec = nullptr;
  std::vector<uint64_t> rmc_key;
  // reptation scratch for rmc_nr reptiles: move draws and per-move sums
  int rmc_nr = 0;
  int32_t* d_rmv = nullptr;
  double* d_rsc = nullptr;
  int32_t* d_rslot = nullptr;  // [nr] evaluation slot of each reptile, [mdelta] counts after it
};

// Queue one batched evaluation (k_pairs, k_gjm or k_gjm_part and k_fgrad, or
// k_det, then k_cross, k_final) on st.
// delta is the time step the drift cutoff of kinetic uses (reptation.f:2755);
// precision QMC_PREC_MIXED selects the mixed-precision pair kernel; nact (a
// device pointer, or null) limits the evaluation to the first *nact walkers.
int launch_eval(qmc_ctx* c, int nw, const double* d_x, double* d_props, double* d_g, double* d_h,
                double* d_g2, int32_t* d_status, hipStream_t st, double delta,
                int precision = QMC_PREC_FP64, const int* nact = nullptr);
// log|psi| alone of nw configurations (d_ltf[w] = ltf, d_status[w] bit 2 =
// singular): the pair kernel's Jastrow pass and the determinant's pivots,
// without F, the drift, the Laplacian or the potential.  d_props receives
// nothing used; the matrix-core determinant path only.
int launch_value(qmc_ctx* c, int nw, const double* d_x, double* d_ltf, int32_t* d_status, double* d_props,
                 hipStream_t st, int precision);
void free_vmc(qmc_ctx* c);
void free_vmc_graph(qmc_ctx* c);
// Workspace of the evaluation kernels for nw walkers (allocates when it grows).
int ensure_ws(qmc_ctx* c, int nw);
int ensure_vmc_scratch(qmc_ctx* c, int nw);
// One VMC step for every walker of w (proposal, evaluation, acceptance;
// block sums when w->blk is set), queued on st.
int vmc_step(qmc_ctx* c, const qmc_walkers* w, hipStream_t st, double delta, double var0, double var1,
             double mage, int bm_order, double* tp, double* ta);
void free_rmc(qmc_ctx* c);


