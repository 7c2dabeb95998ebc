// qmc_cross.hip -- backflow Laplacian cross term on the FP64 matrix cores.
//
// For every spin s and particle type t with backflow between them,
//   h_t += sum_{ip in t, idim} sum_{j,k} Re[(qa F)(qa F)]   (reptation.f:10241-10259)
// which is the symmetric contraction sum_{r,q} M[r][q] S_t[r][q] over rows
// r = (j, a) of the spin's particles (R = 2m <= 128 rows) with
//   M[(j,a),(k,b)] = Re(F_a(k,j) F_b(j,k))            (symmetric),
//   S_t = V_t V_t^T,  V_t[(j,a),(ip,idim)] = qa(a,idim,j,ip),
// qa being the backflow Jacobian: 1 + Q_j on the particle's own block and
// -(eta delta + eta' r u u) for a pair (reptation.f:10023-10058).
//
// One workgroup of 8 wavefronts per (walker, spin).  S_t is a 128 x 128
// SYRK built from 16 x 16 tiles with v_mfma_f64_16x16x4f64: the 36 tiles of
// the upper triangle go 9 to each pair of wavefronts, block rows g and 7-g
// (8-g and g+1 tiles), split 5 + 4 between the pair.  V_t is staged in LDS 16
// particles (32 columns) at a time, column-major so each MFMA operand is one
// 8-byte LDS read per lane.  M is formed from F (L2) tile by tile when S_t is
// complete, once for both particle types, so only the S accumulators stay in
// registers through the chunks and F is read once per (walker, spin); it is
// prefetched into L2 while the last chunk's products run.
#include "qmc_dev.h"

typedef double xc_d4 __attribute__((ext_vector_type(4)));

// Tile i (0..8) of the group for block row g: row g then row 7-g.
__host__ __device__ constexpr int xc_tj(int g, int i) { return i < 8 - g ? g : 7 - g; }
__host__ __device__ constexpr int xc_tk(int g, int i) { return i < 8 - g ? g + i : (7 - g) + (i - (8 - g)); }
// wavefront half h takes tiles [xc_i0(h), xc_i0(h) + xc_nt(h))
__host__ __device__ constexpr int xc_i0(int h) { return h ? 5 : 0; }
__host__ __device__ constexpr int xc_nt(int h) { return h ? 4 : 5; }
__host__ __device__ constexpr bool xc_needs(int g, int h, int b) {
  for (int q = 0; q < xc_nt(h); q++)
    if (xc_tj(g, xc_i0(h) + q) == b || xc_tk(g, xc_i0(h) + q) == b) return true;
  return false;
}

#define XC_TPW 5  // most tiles per wavefront

// Empty statement that takes the accumulator in registers: the MFMAs that
// produce it are issued before this point.
__device__ __forceinline__ void xc_pin(xc_d4& a) { __asm__ volatile("" : "+v"(a)); }

// Sums over the wavefront's tiles of M o S_0 and M o S_1, M formed from F
// here once for both particle types (weight 2 off the diagonal).  The
// accumulator of v_mfma_f64_16x16x4f64 holds, in lane l and register v,
// element (row l/16 + 4v, column l%16) of the 16 x 16 tile.
template <int G, int H>
__device__ __forceinline__ void xc_contract(const xc_d4 (&acc)[2][XC_TPW], const zd* __restrict__ Fg, int m,
                                            int lane_in, int nT, double& t0, double& t1) {
  const int R = 2 * m;
  // the addresses of F are formed here, not hoisted out of the chunk loops
  // (where they would have to be spilled)
  const zd* Fq = Fg;
  int lane = lane_in;
  __asm__ volatile("" : "+s"(Fq), "+v"(lane));
  t0 = t1 = 0.0;
#pragma unroll
  for (int q = 0; q < xc_nt(H); q++) {
    const int TJ = xc_tj(G, xc_i0(H) + q), TK = xc_tk(G, xc_i0(H) + q);
    if (TK < nT) {
      const int sc = 16 * TK + (lane & 15);
      const int ks = sc >> 1, bs = sc & 1;
      double u0 = 0.0, u1 = 0.0;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = 16 * TJ + (lane >> 4) + 4 * v;
        if (r < R && sc < R) {
          const int jr = r >> 1, ar = r & 1;
          const zd f1 = Fq[(ar * m + ks) * m + jr];
          const zd f2 = Fq[(bs * m + jr) * m + ks];
          const double mv = f1.x * f2.x - f1.y * f2.y;
          u0 = fma(mv, acc[0][q][v], u0);
          u1 = fma(mv, acc[1][q][v], u1);
        }
      }
      t0 += TJ == TK ? u0 : 2.0 * u0;
      t1 += TJ == TK ? u1 : 2.0 * u1;
    }
    // one tile's F values in flight at a time
    __asm__ volatile("" ::: "memory");
  }
}

// Entry e (row j = e % 64, particle pc = e / 64) of the V chunk of
// particles c0..c0+cn-1: rows j beyond m and particles beyond cn are zero;
// the 2 x 2 block of (j, ip) is symmetric.  Consecutive threads take
// consecutive rows, so the column-major stores spread over the LDS banks.
__device__ __forceinline__ void xc_entry(const DevSys& s, const double* __restrict__ Qd,
                                         const double* __restrict__ xs, double* __restrict__ Vt, TabRef tb,
                                         int e, int c0, int cn, int m, int p0) {
  const int j = e & 63, pc = e >> 6;
  double v11 = 0, v12 = 0, v22 = 0;
  const int ip = c0 + pc;
  if (j < m && pc < cn) {
    const int jp = p0 + j;
    if (jp == ip) {
      v11 = 1.0 + Qd[3 * j];
      v12 = Qd[3 * j + 1];
      v22 = 1.0 + Qd[3 * j + 2];
    } else if (tb.row0 >= 0) {
      const double2 xj2 = ld2(xs, jp), xi2 = ld2(xs, ip);
      PairGeo pg = pair_geo(s, xj2.x, xj2.y, xi2.x, xi2.y);
      double eta, deta, ddeta;
      spline3(s, tb, pg.ind, pg.rem, eta, deta, ddeta);
      const double cb = deta * pg.byr;
      v11 = -(eta + cb * pg.ux * pg.ux);
      v12 = -(cb * pg.ux * pg.uy);
      v22 = -(eta + cb * pg.uy * pg.uy);
    }
  }
  // rows 2j, 2j+1 of each of the two columns: one 16-byte store each
  // (XC_RS is even, so both are 16-byte aligned)
  *reinterpret_cast<double2*>(Vt + (2 * pc) * XC_RS + 2 * j) = make_double2(v11, v12);
  *reinterpret_cast<double2*>(Vt + (2 * pc + 1) * XC_RS + 2 * j) = make_double2(v12, v22);
}

// Destination of the L2 prefetch of F: one wavefront's 64 x 16 bytes, never read.
__shared__ __attribute__((aligned(16))) unsigned xc_sink[256];

// Per-wavefront body for tile group (G, H).  The V chunks (16 particles of
// one type at a time, over the types with backflow to this spin) are
// double-buffered: while the matrix cores take this chunk's products the
// vector units build the next chunk into the other buffer (two entries per
// thread, placed between the MFMA steps), so a chunk costs one barrier.
template <int G, int H>
__device__ __forceinline__ void xc_body(const DevSys& s, double* __restrict__ dsc, const double* __restrict__ Qd,
                                        const zd* __restrict__ Fg, const double* __restrict__ xs,
                                        double* __restrict__ Vt, double* red, int sp, int m, int p0, int tid) {
  const int lane = tid & 63;
  const int nT = (2 * m + 15) >> 4;
  // S_t accumulators of both particle types (zero for a type without
  // backflow to this spin), contracted with M in one pass over F at the end
  xc_d4 acc[2][XC_TPW];
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int q = 0; q < XC_TPW; q++) acc[t][q] = xc_d4{0.0, 0.0, 0.0, 0.0};
  bool act[2];
#pragma unroll
  for (int t = 0; t < 2; t++)
    act[t] = t < s.ntypes && (sp ? pick2(s.iub_flag[1], t) : pick2(s.iub_flag[0], t)) && pick2(s.np, t) > 0;
  const int ft = act[0] ? 0 : (act[1] ? 1 : -1);
  if (ft >= 0) {
    const int c0 = pick2(s.ipf, ft), cn = min(XC_KC / 2, pick2(s.np, ft));
    const TabRef tb = tabref<TK_B>(s, tpair(sp, ft));
    for (int e = tid; e < 64 * (XC_KC / 2); e += XC_THREADS) xc_entry(s, Qd, xs, Vt, tb, e, c0, cn, m, p0);
  }
  __syncthreads();
  int cur = 0;
#pragma unroll
  for (int t = 0; t < 2; t++) {
    if (!act[t]) continue;
    const int ia = pick2(s.ipf, t), ib = ia + pick2(s.np, t);
    for (int c0 = ia; c0 < ib; c0 += XC_KC / 2) {
      // the next chunk: the rest of this type, then the other type
      int nt = t, nc0 = c0 + XC_KC / 2;
      if (nc0 >= ib) {
        nt = (t == 0 && act[1]) ? 1 : -1;
        nc0 = nt >= 0 ? pick2(s.ipf, 1) : 0;
      }
      const int ncn = nt >= 0 ? min(XC_KC / 2, pick2(s.ipf, nt) + pick2(s.np, nt) - nc0) : 0;
      const TabRef ntb = nt >= 0 ? tabref<TK_B>(s, tpair(sp, nt)) : TabRef{-1, 0};
      double* Vn = Vt + (cur ^ 1) * (XC_KC * XC_RS);
      const double* base = Vt + cur * (XC_KC * XC_RS) + (lane >> 4) * XC_RS + (lane & 15);
      // particles in this chunk: its columns 2 cn and beyond are zero, and
      // a step whose four columns all are is skipped
      const int cn = min(XC_KC / 2, ib - c0);
#pragma unroll
      for (int ks = 0; ks < XC_KC / 4; ks++) {
        if (2 * ks < cn) {
          // rows beyond 2m are zero in LDS, so tiles past the last block row
          // add zeros and need no branch here
          double blk[8];
#pragma unroll
          for (int b = 0; b < 8; b++)
            if (xc_needs(G, H, b)) blk[b] = base[ks * 4 * XC_RS + 16 * b];
#pragma unroll
          for (int q = 0; q < xc_nt(H); q++) {
            const int TJ = xc_tj(G, xc_i0(H) + q), TK = xc_tk(G, xc_i0(H) + q);
            acc[t][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(blk[TJ], blk[TK], acc[t][q], 0, 0, 0);
          }
#pragma unroll
          for (int q = 0; q < xc_nt(H); q++) xc_pin(acc[t][q]);
        }
        if (nt >= 0 && (ks == 1 || ks == 5))
          xc_entry(s, Qd, xs, Vn, ntb, tid + (ks == 5 ? XC_THREADS : 0), nc0, ncn, m, p0);
        // during the last chunk, F of this (walker, spin) is fetched into
        // the L2 cache (16 bytes a lane, into an LDS sink nothing reads), a
        // quarter at every other step, so the contraction below reads it
        // from there instead of waiting on HBM tile by tile
        if (nt < 0 && (ks & 1)) {
          const int nz = 2 * m * m;
          for (int i = tid + (ks >> 1) * XC_THREADS; i < nz; i += 4 * XC_THREADS)
            __builtin_amdgcn_global_load_lds((const void*)(Fg + i), (void __attribute__((address_space(3)))*)xc_sink,
                                             16, 0, 0);
        }
        // the next step's operands are read after this step's MFMAs issue
        __asm__ volatile("" ::: "memory");
      }
      __syncthreads();
      cur ^= 1;
    }
  }
  double u0, u1;
  xc_contract<G, H>(acc, Fg, m, lane, nT, u0, u1);
  u0 = block_sum<XC_THREADS>(u0, red);
  u1 = block_sum<XC_THREADS>(u1, red);
  if (tid == 0) {
    dsc[DS_X0] = u0;
    dsc[DS_X1] = u1;
  }
}

// LDS: xs[2N], then two Vt[XC_KC][XC_RS] chunk buffers (column = (particle,
// idim) of the chunk, row = (j, a)).
__global__ __launch_bounds__(XC_THREADS, 4) void k_cross(DevSys s, WsLayout L, const double* __restrict__ x,
                                                         double* __restrict__ ws,
                                                         const double* __restrict__ Fbuf) {
  extern __shared__ __attribute__((aligned(16))) double sm[];
  __shared__ double red[XC_THREADS / 64];
  // walkers and spins in the reverse of the order in which k_gjm wrote their
  // F, so the first workgroups find theirs still in the memory-side cache
  const int w = gridDim.x - 1 - blockIdx.x, sp = gridDim.y - 1 - blockIdx.y;
  if (s.nact && w >= *s.nact) return;
  const int m = pick2(s.np, sp);
  if (m == 0) return;
  const int N = s.nptot, p0 = pick2(s.ipf, sp);
  const int tid = threadIdx.x;
  const int w8 = __builtin_amdgcn_readfirstlane(tid >> 6);
  double* wsw = ws + (int64_t)w * L.per_walker;
  double* dsc = wsw + pick2(L.off_det, sp);
  if (dsc[DS_SING] != 0.0) return;
  const double* Qd = dsc + DS_N + 2 * N;
  const zd* Fg = (const zd*)(Fbuf + (int64_t)w * L.f_per_walker + pick2(L.f_off, sp));
  double* xs = sm;
  double* Vt = sm + 2 * N;
  const double* xw = x + (int64_t)w * 2 * N;
  for (int k = tid; k < 2 * N; k += XC_THREADS) xs[k] = xw[k];
  __syncthreads();
  switch (w8) {
    case 0: xc_body<0, 0>(s, dsc, Qd, Fg, xs, Vt, red, sp, m, p0, tid); break;
    case 1: xc_body<1, 0>(s, dsc, Qd, Fg, xs, Vt, red, sp, m, p0, tid); break;
    case 2: xc_body<2, 0>(s, dsc, Qd, Fg, xs, Vt, red, sp, m, p0, tid); break;
    case 3: xc_body<3, 0>(s, dsc, Qd, Fg, xs, Vt, red, sp, m, p0, tid); break;
    case 4: xc_body<0, 1>(s, dsc, Qd, Fg, xs, Vt, red, sp, m, p0, tid); break;
    case 5: xc_body<1, 1>(s, dsc, Qd, Fg, xs, Vt, red, sp, m, p0, tid); break;
    case 6: xc_body<2, 1>(s, dsc, Qd, Fg, xs, Vt, red, sp, m, p0, tid); break;
    default: xc_body<3, 1>(s, dsc, Qd, Fg, xs, Vt, red, sp, m, p0, tid); break;
  }
}
