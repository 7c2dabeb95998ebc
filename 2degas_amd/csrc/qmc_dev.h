// qmc_dev.h -- device-side definitions shared by the walker-evaluation kernels.
//
// One DevSys value (kernel argument) carries the system description and the
// device pointers of the deck tables; the per-walker workspace layout is
// described by WsLayout.  All arithmetic here is FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define QMC_MAXTAB 20
#define QMC_MAX_ORB 64
#define QMC_BLOCK 256

// Type pair index of particles of types ti, tj (0: 1-1, 1: 1-2, 2: 2-2).
__host__ __device__ inline int tpair(int ti, int tj) { return ti + tj; }

struct DevSys {
  int ntypes, nptot, npnorm;
  int np[2], ipf[2];
  double hbs2m[2], el[2], eli[2], drt, drti, drti2;
  // spline tables: rows of 4 coefficients, table t row i at tab[t*tab_ld + i]
  const double4* tab;
  int tab_ld;
  int ngrid[QMC_MAXTAB];
  int iexp[QMC_MAXTAB];
  // table (0-based, -1 = none) per type pair for u2, u3, backflow, v2
  int tu2[3], tu3[3], tub[3], tv2[3];
  int has_u2, has_u3, has_b, has_v2;
  int iub_flag[2][2];      // determinant of spin s couples particle type t
  double v2value[3];
  double tailc[2], tailc_tot;  // pair-potential tail constants
  // k space
  int nk, nrhok, irhok[2];
  const double* kvec;      // [nk][2]
  int nke[3];              // nk_ewald of the v2 table of each type pair
  const double* vk[3];     // [nke]
  double vk0[3];
  int ntheta;
  double v0, adrift, delta;
  // property slots (0-based)
  int n_props, jetot, jltf, jpot0, jkin0, je2, jgg, jefp, jrhok;
  int jpot[2], jkin[2], jun[2];
  // orbitals
  const double* gvec;      // [ng][2]
  int ng, has_det;
};

// Per-walker workspace (doubles) and per-spin determinant buffers.
struct WsLayout {
  int64_t per_walker;      // doubles per walker in ws
  int off_gj, off_q, off_sc;   // jastrow gradient [N][2], quasi-coords [N][2], scalars
  int off_det[2];          // per spin: [DET_SC] scalars + g_det [N][2] + Qd [m][3]
  int64_t f_per_walker;    // doubles per walker in the F buffer (both spins)
  int64_t f_off[2];        // offset of spin s in the F buffer
  int64_t a_per_walker;    // doubles per walker in the A-copy buffer
  int64_t a_off[2];
};

// scalar slots in off_sc
enum { SC_LTF = 0, SC_H0 = 1, SC_H1 = 2, SC_P0 = 3, SC_P1 = 4, SC_PT = 5, SC_N = 8 };
// scalar slots in off_det[s]
enum { DS_LOGDET = 0, DS_H0 = 1, DS_H1 = 2, DS_G2 = 3, DS_X0 = 4, DS_X1 = 5, DS_SING = 6, DS_N = 8 };

struct zd {
  double x, y;
};
__device__ __forceinline__ zd zmk(double a, double b) { return zd{a, b}; }
__device__ __forceinline__ zd zmul(zd a, zd b) { return zd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ zd zadd(zd a, zd b) { return zd{a.x + b.x, a.y + b.y}; }
// c + a*b
__device__ __forceinline__ zd zfma(zd a, zd b, zd c) {
  return zd{fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}
__device__ __forceinline__ double cabs1(zd a) { return fabs(a.x) + fabs(a.y); }
__device__ __forceinline__ zd zrcp(zd a) {
  double d = a.x * a.x + a.y * a.y;
  return zd{a.x / d, -a.y / d};
}

// Cubic spline on the uniform grid, value and two derivatives
// (getf, reptation.f:4114-4125).
__device__ __forceinline__ void spline3(const DevSys& s, int t, int ind, double rem, double& f,
                                        double& df, double& ddf) {
  int i = ind < s.ngrid[t] ? (ind > 0 ? ind : 0) : s.ngrid[t];
  double4 c = s.tab[(int64_t)t * s.tab_ld + i];
  double rm = rem * s.drti;
  f = c.x + rm * (c.y + rm * (c.z + rm * c.w));
  df = (c.y + rm * (2.0 * c.z + rm * (3.0 * c.w))) * s.drti;
  ddf = (c.z + rm * (3.0 * c.w)) * s.drti2;
}

// Value only, times (1/r)^iexp (getf_0, reptation.f:4103-4112).
__device__ __forceinline__ double spline0(const DevSys& s, int t, int ind, double rem, double byr) {
  int i = ind < s.ngrid[t] ? (ind > 0 ? ind : 0) : s.ngrid[t];
  double4 c = s.tab[(int64_t)t * s.tab_ld + i];
  double rm = rem * s.drti;
  double f = c.x + rm * (c.y + rm * (c.z + rm * c.w));
  int e = s.iexp[t];
  double p = 1.0;
  int n = e < 0 ? -e : e;
  for (int k = 0; k < n; k++) p *= byr;
  return e < 0 ? f / p : f * p;
}

// Minimum-image pair geometry (distances, reptation.f:4138-4151):
// u = x_i - x_j folded with nint, r, grid index and remainder, 1/r.
struct PairGeo {
  double ux, uy, r, rem, byr;
  int ind;
};
__device__ __forceinline__ PairGeo pair_geo(const DevSys& s, double xi, double yi, double xj,
                                            double yj) {
  PairGeo p;
  double dx = xi - xj, dy = yi - yj;
  dx = dx - s.el[0] * round(s.eli[0] * dx);
  dy = dy - s.el[1] * round(s.eli[1] * dy);
  p.ux = dx;
  p.uy = dy;
  p.r = sqrt(dx * dx + dy * dy);
  p.ind = (int)(p.r * s.drti);
  p.rem = p.r - p.ind * s.drt;
  p.byr = 1.0 / p.r;
  return p;
}

__device__ __forceinline__ int ptype(const DevSys& s, int i) { return i >= s.ipf[1] && s.ntypes > 1 ? 1 : 0; }

// Block-wide sum of a double over QMC_BLOCK threads; red must hold
// QMC_BLOCK/64 doubles.  All threads get the result.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < QMC_BLOCK / 64; k++) t += red[k];
  return t;
}
