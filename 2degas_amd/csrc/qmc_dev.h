// qmc_dev.h -- device-side definitions shared by the walker-evaluation kernels.
//
// One DevSys value (kernel argument) carries the system description and the
// device pointers of the deck tables; the per-walker workspace layout is
// described by WsLayout.  Arithmetic is FP64, except the FP32 table
// evaluations the mixed-precision pair kernel selects (spline3m, spline0<true>).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define QMC_MAXTAB 20
#define QMC_MAX_ORB 64
#define QMC_BLOCK 256
// k_cross: threads, V columns staged per chunk, LDS row stride (doubles)
#define XC_THREADS 512
#define XC_KC 32
#define XC_RS 144  // 128 rows + 16: the four 16-lane groups of an MFMA operand read disjoint banks

// Type pair index of particles of types ti, tj (0: 1-1, 1: 1-2, 2: 2-2).
__host__ __device__ inline int tpair(int ti, int tj) { return ti + tj; }

// Per type pair: the first row and grid size of its u2, u3, backflow and
// potential tables (row -1: no table), the potential's 1/r exponent and its
// weight.  A copy lives in device memory so that a loop over type pairs
// reads it with scalar loads instead of keeping every pair's values live.
struct __attribute__((aligned(16))) TpInfo {
  int row[4], ng[4];
  int vexp, pad;
  double v2value;
};

struct DevSys {
  int ntypes, nptot, npnorm;
  int np[2], ipf[2];
  double hbs2m[2], el[2], eli[2], drt, drti, drti2;
  // spline tables: rows of 4 coefficients, table t row i at tab[t*tab_ld + i];
  // tab32 holds the same rows rounded to FP32 (the mixed-precision path)
  const double4* tab;
  const float4* tab32;
  int tab_ld;
  int ngrid[QMC_MAXTAB];
  int iexp[QMC_MAXTAB];
  // table (0-based, -1 = none) per type pair for u2, u3, backflow, v2
  int tu2[3], tu3[3], tub[3], tv2[3];
  int has_u2, has_u3, has_b, has_v2;
  // per table kind (TK_U2, TK_U3, TK_B, TK_V) and type pair: first row of
  // the table in `tab` and its ngrid (-1 row = no table); iexp of v tables
  int trow[4][3], tng[4][3], vexp[3];
  int iub_flag[2][2];      // determinant of spin s couples particle type t
  double v2value[3];
  const TpInfo* tpi;       // [3], the same tables per type pair
  double tailc[2], tailc_tot;  // pair-potential tail constants
  // k space
  int nk, nrhok, irhok[2];
  const double* kvec;      // [nk][2]
  // factorized phases: k = (kxv[kxi], kyv[kyi]) over few distinct components;
  // nkxv = 0 selects the direct sincos path
  int nkxv, nkyv, kchunk;
  const double* kxv;
  const double* kyv;
  // kprog: kxv and kyv are ascending arithmetic progressions with steps dkx,
  // dky (lattice k-vectors), so runs of phases follow from one by products
  int kprog;
  double dkx, dky;
  const int* kxy;          // [nkslot] factorized rho_k slots: kxi | kyi << 8 | (k + 1) << 16
  int nkslot;
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
  // walkers evaluated: the first *nact of the batch when not null (the
  // reptation sweep evaluates only the reptiles that still move a bead)
  const int* nact;
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

// A constant of qsincos; with SREG it passes through an empty asm
// statement bound to a scalar register pair, which keeps it in SGPRs and
// stops it being hoisted out of a loop into vector registers (the same value
// either way).
template <bool SREG>
__device__ __forceinline__ double qk(double c) {
  if (SREG) __asm__ volatile("" : "+s"(c));
  return c;
}

// sin and cos of x for |x| < 2^30 (absolute error below 1.4e-16).  Cody-Waite
// reduction r = x - n pi/2 with pi/2 split over three doubles (the first
// product is exact inside the FMA), then the Taylor series of sin r and cos r
// on |r| <= pi/4 through r^17 and r^18 (coefficients 1/k!, truncation below
// 1e-19), turned by the quadrant n mod 4.  Uses a handful of registers, where
// the library routine's wide-argument path needs about a hundred.
template <bool SREG>
__device__ __forceinline__ void qsincos_t(double x, double* sn, double* cs) {
  const double n = rint(x * qk<SREG>(0x1.45f306dc9c883p-1));
  double r = fma(-n, qk<SREG>(0x1.921fb54442d18p+0), x);
  r = fma(-n, qk<SREG>(0x1.1a62633145c07p-54), r);
  r = fma(-n, qk<SREG>(-0x1.f1976b7ed8fbcp-110), r);
  const double r2 = r * r;
  double ps = qk<SREG>(1.0 / 355687428096000.0);
  ps = fma(ps, r2, qk<SREG>(-1.0 / 1307674368000.0));
  ps = fma(ps, r2, qk<SREG>(1.0 / 6227020800.0));
  ps = fma(ps, r2, qk<SREG>(-1.0 / 39916800.0));
  ps = fma(ps, r2, qk<SREG>(1.0 / 362880.0));
  ps = fma(ps, r2, qk<SREG>(-1.0 / 5040.0));
  ps = fma(ps, r2, qk<SREG>(1.0 / 120.0));
  ps = fma(ps, r2, qk<SREG>(-1.0 / 6.0));
  const double sr = fma(r * r2, ps, r);
  double pc = qk<SREG>(-1.0 / 6402373705728000.0);
  pc = fma(pc, r2, qk<SREG>(1.0 / 20922789888000.0));
  pc = fma(pc, r2, qk<SREG>(-1.0 / 87178291200.0));
  pc = fma(pc, r2, qk<SREG>(1.0 / 479001600.0));
  pc = fma(pc, r2, qk<SREG>(-1.0 / 3628800.0));
  pc = fma(pc, r2, qk<SREG>(1.0 / 40320.0));
  pc = fma(pc, r2, qk<SREG>(-1.0 / 720.0));
  pc = fma(pc, r2, qk<SREG>(1.0 / 24.0));
  pc = fma(pc, r2, qk<SREG>(-0.5));
  const double cr = fma(pc, r2, 1.0);
  const int q = (int)((long long)n & 3);
  const double sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
  *sn = (q & 2) ? -sv : sv;
  *cs = ((q + 1) & 2) ? -cv : cv;
}

__device__ __forceinline__ void qsincos(double x, double* sn, double* cs) { qsincos_t<false>(x, sn, cs); }
// qsincos with its coefficients kept in scalar registers, for loops whose
// vector registers are all taken
__device__ __forceinline__ void qsincos_sreg(double x, double* sn, double* cs) { qsincos_t<true>(x, sn, cs); }

// x <- a x mod 2^48 with a = 34522712143931 (the limbs 502, 1521, 4071, 2107
// of reptation.f:5349-5354); u = x / 2^48 exactly.  Unsigned 64-bit products
// wrap mod 2^64, a multiple of 2^48, so the low 48 bits are exact.
__device__ __forceinline__ double rannyu(uint64_t& x) {
  const uint64_t a = 34522712143931ULL;
  x = (a * x) & ((1ULL << 48) - 1ULL);
  return (double)x * 0x1p-48;
}

// a^n mod 2^48 for rannyu's multiplier: the stream n draws ahead is a^n x.
__device__ __forceinline__ uint64_t lcg_pow(uint64_t n) {
  uint64_t p = 1, q = 34522712143931ULL;
  for (; n; n >>= 1) {
    if (n & 1) p = (p * q) & ((1ULL << 48) - 1ULL);
    q = (q * q) & ((1ULL << 48) - 1ULL);
  }
  return p;
}

enum { TK_U2 = 0, TK_U3 = 1, TK_B = 2, TK_V = 3 };

// Element i of a 2- or 3-entry member array without a dynamic index (a
// dynamic index into the kernel-argument struct would copy it to scratch).
template <typename T>
__device__ __forceinline__ T pick2(const T (&a)[2], int i) { return i ? a[1] : a[0]; }
template <typename T>
__device__ __forceinline__ T pick3(const T (&a)[3], int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

struct TabRef {
  int row0, ng;
};
template <int KIND>
__device__ __forceinline__ TabRef tabref(const DevSys& s, int tp) {
  return TabRef{pick3(s.trow[KIND], tp), pick3(s.tng[KIND], tp)};
}

// Cubic spline on the uniform grid, value and two derivatives
// (getf, reptation.f:4114-4125).
__device__ __forceinline__ void spline3(const DevSys& s, TabRef t, int ind, double rem, double& f,
                                        double& df, double& ddf) {
  int i = ind < t.ng ? (ind > 0 ? ind : 0) : t.ng;
  double4 c = s.tab[t.row0 + i];
  double rm = rem * s.drti;
  f = c.x + rm * (c.y + rm * (c.z + rm * c.w));
  df = (c.y + rm * (2.0 * c.z + rm * (3.0 * c.w))) * s.drti;
  ddf = (c.z + rm * (3.0 * c.w)) * s.drti2;
}

// spline3 on the FP32 copy of the tables when MIX: the cubic and its
// derivatives in FP32 from the FP64 grid remainder, returned widened to FP64.
template <bool MIX>
__device__ __forceinline__ void spline3m(const DevSys& s, TabRef t, int ind, double rem, double& f,
                                         double& df, double& ddf) {
  if (!MIX) {
    spline3(s, t, ind, rem, f, df, ddf);
    return;
  }
  int i = ind < t.ng ? (ind > 0 ? ind : 0) : t.ng;
  const float4 c = s.tab32[t.row0 + i];
  const float rm = (float)(rem * s.drti);
  const float di = (float)s.drti;
  f = (double)(c.x + rm * (c.y + rm * (c.z + rm * c.w)));
  df = (double)((c.y + rm * (2.0f * c.z + rm * (3.0f * c.w))) * di);
  ddf = (double)((c.z + rm * (3.0f * c.w)) * (float)s.drti2);
}

// Value only, times (1/r)^iexp (getf_0, reptation.f:4103-4112).
template <bool MIX = false>
__device__ __forceinline__ double spline0(const DevSys& s, TabRef t, int e, int ind, double rem,
                                         double byr) {
  int i = ind < t.ng ? (ind > 0 ? ind : 0) : t.ng;
  double f;
  if (MIX) {
    const float4 c = s.tab32[t.row0 + i];
    const float rm = (float)(rem * s.drti);
    f = (double)(c.x + rm * (c.y + rm * (c.z + rm * c.w)));
  } else {
    const double4 c = s.tab[t.row0 + i];
    const double rm = rem * s.drti;
    f = c.x + rm * (c.y + rm * (c.z + rm * c.w));
  }
  double p = 1.0;
  int n = e < 0 ? -e : e;
  for (int k = 0; k < n; k++) p *= byr;
  return e < 0 ? f / p : f * p;
}

// r = sqrt(r2) and 1/r for r2 > 0: the hardware reciprocal square root
// (about 23 correct bits) refined by two Newton steps y <- y + y (1/2 - r2 y^2 / 2),
// which leave y within a few ulp of 1/sqrt(r2); r = r2 y.  Some 10
// instructions where the correctly rounded sqrt and divide take over 30.
__device__ __forceinline__ void qrsqrt(double r2, double& r, double& byr) {
  double y = __builtin_amdgcn_rsq(r2);
  const double h = 0.5 * r2;
  double e = fma(-(h * y), y, 0.5);
  y = fma(y, e, y);
  e = fma(-(h * y), y, 0.5);
  y = fma(y, e, y);
  byr = y;
  r = r2 * y;
}

// Minimum-image pair geometry (distances, reptation.f:4138-4151):
// u = x_i - x_j folded with nint, r, grid index and remainder, 1/r.
struct PairGeo {
  double ux, uy, r, rem, byr;
  int ind;
};
// round(t), halves away from zero (Fortran's nint), as trunc(t + copysign(h, t))
// with h the double just below 1/2: equal to round() for every double (ties,
// their neighbours and random bit patterns checked on the CPU), in two FP64
// operations and a bit insert where round() takes six.
__device__ __forceinline__ double qround(double t) { return trunc(t + copysign(0.49999999999999994, t)); }

// Entry i of an LDS array of (x, y) pairs as one 16-byte read (the array
// starts 16-byte aligned): two 8-byte reads of neighbouring lanes' pairs
// would each stride the banks at 16 bytes.
__device__ __forceinline__ double2 ld2(const double* a, int i) { return *reinterpret_cast<const double2*>(a + 2 * i); }

__device__ __forceinline__ PairGeo pair_geo(const DevSys& s, double xi, double yi, double xj,
                                            double yj) {
  PairGeo p;
  double dx = xi - xj, dy = yi - yj;
  dx = dx - s.el[0] * qround(s.eli[0] * dx);
  dy = dy - s.el[1] * qround(s.eli[1] * dy);
  p.ux = dx;
  p.uy = dy;
  qrsqrt(dx * dx + dy * dy, p.r, p.byr);
  p.ind = (int)(p.r * s.drti);
  p.rem = p.r - p.ind * s.drt;
  return p;
}

__device__ __forceinline__ int ptype(const DevSys& s, int i) { return i >= s.ipf[1] && s.ntypes > 1 ? 1 : 0; }

// k_fgrad's schedule for one spin: nA = 32 steps of the spin's own block
// (weight 3: both ends take drift terms) then nB = 64 steps across to the
// other type (weight 2), split into four runs of equal weight, one per
// wavefront.  A wavefront's LDS partial sums hold the backflow sums qb[5][64]
// of the spin's particles, then zs[6][64] (drift and Laplacian terms of the
// spin's particles) when its run has own-block steps and zo[6][64] (those of
// the other type) when it has steps across.  Scalar functions of (nA, nB, w)
// rather than a table, which a varying w would index in scratch memory.
#define FG_NZ 6
#define FG_NQ 5
// first step of wavefront w's run (w = 4: the end)
__host__ __device__ inline int fg_u(int nA, int nB, int w) {
  if (w >= 4) return nA + nB;
  const int C = 3 * nA + 2 * nB, T = (w * C) / 4;
  const int u = T <= 3 * nA ? (T + 2) / 3 : nA + (T - 3 * nA + 1) / 2;
  return u < nA + nB ? u : nA + nB;
}
__host__ __device__ inline bool fg_hasA(int nA, int nB, int w) {
  return fg_u(nA, nB, w + 1) > fg_u(nA, nB, w) && fg_u(nA, nB, w) < nA;
}
__host__ __device__ inline bool fg_hasB(int nA, int nB, int w) {
  return fg_u(nA, nB, w + 1) > fg_u(nA, nB, w) && fg_u(nA, nB, w + 1) > nA;
}
// start of wavefront w's partial sums (w = 4: their total), in doubles
__host__ __device__ inline int fg_off(int nA, int nB, int w) {
  int o = 0;
  for (int v = 0; v < w; v++)
    o += (fg_u(nA, nB, v + 1) > fg_u(nA, nB, v) ? FG_NQ * 64 : 0) + (fg_hasA(nA, nB, v) ? FG_NZ * 64 : 0) +
         (fg_hasB(nA, nB, v) ? FG_NZ * 64 : 0);
  return o;
}

// Maximum over the wavefront of a 32-bit key: DPP swaps inside quads, row
// rotations by 4 and 8 inside each row of 16 lanes, then the four row maxima.
// Every lane must be active.
__device__ __forceinline__ unsigned gj_wave_max(unsigned x) {
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x124, 0xF, 0xF, false));
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xF, 0xF, false));
  const unsigned r0 = __builtin_amdgcn_readlane((int)x, 0);
  const unsigned r1 = __builtin_amdgcn_readlane((int)x, 16);
  const unsigned r2 = __builtin_amdgcn_readlane((int)x, 32);
  const unsigned r3 = __builtin_amdgcn_readlane((int)x, 48);
  return max(max(r0, r1), max(r2, r3));
}

__device__ __forceinline__ double gj_readlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Block-wide sum of a double over QMC_BLOCK threads; red must hold
// QMC_BLOCK/64 doubles.  All threads get the result.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// Sum over a workgroup of NT threads (red holds NT / 64 doubles).
template <int NT = QMC_BLOCK>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < NT / 64; k++) t += red[k];
  return t;
}
