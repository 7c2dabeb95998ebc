// qmc_kernels.hip -- FP64 walker-evaluation kernels for gfx950.
//
// One local-energy evaluation (compute_properties(1), reptation.f:2661-2727)
// of a batch of walkers is five launches: k_pairs, k_gjm and k_fgrad
// (qmc_gjm.hip, qmc_fgrad.hip: the determinant on the matrix cores), k_cross
// and k_final; with QMC_NO_GJ set, k_det takes the part of k_gjm and k_fgrad:
//   k_pairs   one workgroup per walker: pair geometry, u2 and u3 Jastrow
//             (two_body/three_body/dotg, reptation.f:2777-3046), backflow
//             quasi-coordinates (reptation.f:10023-10058), real-space Ewald
//             (reptation.f:3994-4033), rho_k and k-space Ewald
//             (reptation.f:4224-4243, 4034-4052).
//   k_det     one workgroup per (walker, spin): plane-wave orbital matrix of
//             the quasi-coordinates, LU with LINPACK pivoting and inverse in
//             LDS (zgefa/zgedi, reptation.f:9227-9472), F = A^-1 dA, drift
//             and the diagonal Laplacian terms (reptation.f:10159-10240).
//   k_cross   (qmc_cross.hip) one workgroup per (walker, spin): the backflow
//             Laplacian cross term sum_ip tr(C_ip^2) (reptation.f:10241-10259)
//             as a symmetric contraction of S = V V^T against M = Re(F o F^T).
//   k_final   kinetic and potential energies and the property slots
//             (kinetic, reptation.f:2729-2775; potential tail, 4088-4096).
#include <type_traits>

#include "qmc_dev.h"

// k-vectors per thread and pass in k_pairs
#define KMAXPT 4
#define KRUN 9  // phases of one particle built from one sin/cos pair (odd: lanes' runs start in different LDS banks)

// ---------------------------------------------------------------------------
// k_pairs
// LDS: xs[2N], Gs[2N], red[8], then a region shared by the per-wavefront
// partial sums of the pair passes (4 x 6 x N doubles) and, afterwards, the
// kchunk x (nkxv + nkyv) phases of rho_k.
//
// The pair passes visit every unordered pair once, on a round-robin
// schedule: the particles of type t sit in lanes 0..np_t-1 of a block of 64
// (np_t <= 64), and step d of block pair (I, J) pairs lane l of I with lane
// (l + d) mod 64 of J, so within a step every lane has a distinct partner.
// Within one type d runs over 1..32 (d = 32 on lanes 0..31 only), across the
// two types over 0..63; the steps are split into four contiguous runs, one
// per wavefront.  A pair's contribution to i stays in registers until the
// wavefront's run leaves block I; its contribution to j is added to the
// wavefront's own partial sums in LDS (distinct j within a step, so no two
// lanes touch one address), and the four partial sums are added in a fixed
// order afterwards: the result does not depend on scheduling.
#define PP_NQ 6  // per-particle sums of a pass

// Step u of the schedule: block (particle type) pair and lane offset.
__device__ __forceinline__ void rr_unit(int u, bool two, int t1, int& ti, int& tj, int& d) {
  if (!two) { ti = tj = t1; d = u + 1; }
  else if (u < 32) { ti = tj = 0; d = u + 1; }
  else if (u < 64) { ti = tj = 1; d = u - 31; }
  else { ti = 0; tj = 1; d = u - 64; }
}

// part[q * N + p] += v for the wavefront's partial sums
__device__ __forceinline__ void pp_add(double* part, int N, int p, const double (&v)[PP_NQ], int nq) {
#pragma unroll
  for (int q = 0; q < PP_NQ; q++)
    if (q < nq) part[q * N + p] += v[q];
}

// MIX (the mixed-precision evaluation): the u2, u3 and pair-potential
// tables are read from their FP32 copy and evaluated in FP32 (spline3m), and
// the rho_k phases are stored in LDS as FP32 pairs and multiplied and summed
// over each chunk of particles in FP32, the chunk sums added in FP64; the
// pair geometry, the backflow coordinates, every accumulator and the k-space
// sum stay FP64.
template <bool MIX>
__device__ __forceinline__ void k_pairs_body(const DevSys& s, const WsLayout& L,
                                             const double* __restrict__ x,
                                             double* __restrict__ ws,
                                             double* __restrict__ props) {
  extern __shared__ __attribute__((aligned(16))) double sm[];
  const int w = blockIdx.x;
  if (s.nact && w >= *s.nact) return;
  const int N = s.nptot;
  double* xs = sm;
  double* Gs = sm + 2 * N;
  double* red = sm + 4 * N;
  double* parts = sm + 4 * N + 8;
  const double* xw = x + (int64_t)w * 2 * N;
  const int tid = threadIdx.x, l = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  double* part = parts + wv * PP_NQ * N;
  for (int k = tid; k < 2 * N; k += QMC_BLOCK) xs[k] = xw[k];
  for (int k = tid; k < 4 * PP_NQ * N; k += QMC_BLOCK) parts[k] = 0.0;
  __syncthreads();
  double* wsw = ws + (int64_t)w * L.per_walker;
  double acc_ltf = 0.0, acc_h0 = 0.0, acc_h1 = 0.0, acc_v0 = 0.0, acc_v1 = 0.0;
  // schedule: both types present, or the one type t1
  const bool two = s.ntypes > 1 && s.np[0] > 0 && s.np[1] > 0;
  const int t1 = (s.ntypes > 1 && s.np[0] == 0) ? 1 : 0;
  const int nU = two ? 128 : 32;
  const int u0 = wv * nU / 4, u1 = (wv + 1) * nU / 4;

  // pass 1: u2, the pair term and G of u3, backflow q, pair potential
  {
    double ia[PP_NQ] = {0, 0, 0, 0, 0, 0};  // gx, gy, Gx, Gy, qx, qy of particle i
    int icur = -1;
    for (int u = u0; u < u1; u++) {
      int ti, tj, d;
      rr_unit(u, two, t1, ti, tj, d);
      if (ti != icur) {
        if (icur >= 0 && l < pick2(s.np, icur)) pp_add(part, N, pick2(s.ipf, icur) + l, ia, PP_NQ);
#pragma unroll
        for (int q = 0; q < PP_NQ; q++) ia[q] = 0.0;
        icur = ti;
      }
      const int jl = (l + d) & 63;
      if (l < pick2(s.np, ti) && jl < pick2(s.np, tj) && (ti != tj || d < 32 || l < 32)) {
        const int i = pick2(s.ipf, ti) + l, j = pick2(s.ipf, tj) + jl;
        const TpInfo* T = s.tpi + tpair(ti, tj);
        const TabRef tu2{T->row[TK_U2], T->ng[TK_U2]}, tu3{T->row[TK_U3], T->ng[TK_U3]},
                     tb{T->row[TK_B], T->ng[TK_B]}, tv{T->row[TK_V], T->ng[TK_V]};
        const double2 xi2 = ld2(xs, i), xj2 = ld2(xs, j);
        PairGeo p = pair_geo(s, xi2.x, xi2.y, xj2.x, xj2.y);
        double ja[PP_NQ] = {0, 0, 0, 0, 0, 0};
        double hh = 0.0;
        double f, df, ddf;
        if (tu2.row0 >= 0) {
          spline3m<MIX>(s, tu2, p.ind, p.rem, f, df, ddf);
          acc_ltf += f;
          const double dtr = df * p.byr;
          ja[0] -= dtr * p.ux;
          ja[1] -= dtr * p.uy;
          hh += ddf + dtr;
        }
        if (tu3.row0 >= 0) {
          spline3m<MIX>(s, tu3, p.ind, p.rem, f, df, ddf);
          ja[2] += f * p.ux;
          ja[3] += f * p.uy;
          // (xi r)^2 pair term of the three-body factor (reptation.f:2822-2832)
          const double r = p.r;
          const double tr = f * r;
          const double d3 = 2.0 * f * (f + r * df);
          const double dd3 = 2.0 * (f * f + 4.0 * r * f * df + r * r * (df * df + ddf * f));
          acc_ltf += tr * tr;
          ja[0] -= d3 * p.ux;
          ja[1] -= d3 * p.uy;
          hh += dd3 + d3;
        }
        if (tb.row0 >= 0) {
          spline3(s, tb, p.ind, p.rem, f, df, ddf);
          ja[4] -= f * p.ux;
          ja[5] -= f * p.uy;
        }
        // the pair gives i the opposite of j's vector terms (u -> -u)
#pragma unroll
        for (int q = 0; q < PP_NQ; q++) ia[q] -= ja[q];
        if (ti == 0) acc_h0 += hh; else acc_h1 += hh;
        if (tj == 0) acc_h0 += hh; else acc_h1 += hh;
        if (tv.row0 >= 0) {
          const double vv = 0.5 * spline0<MIX>(s, tv, T->vexp, p.ind, p.rem, p.byr) * T->v2value;
          if (ti == 0) acc_v0 += vv; else acc_v1 += vv;
          if (tj == 0) acc_v0 += vv; else acc_v1 += vv;
        }
        pp_add(part, N, j, ja, PP_NQ);
      }
    }
    if (icur >= 0 && l < pick2(s.np, icur)) pp_add(part, N, pick2(s.ipf, icur) + l, ia, PP_NQ);
  }
  __syncthreads();
  // per-particle sums in a fixed order; Gs for pass 2, q out, g kept
  double gpx = 0.0, gpy = 0.0;  // gradient of particle tid (N <= 128)
  for (int p = tid; p < N; p += QMC_BLOCK) {
    double v[PP_NQ];
#pragma unroll
    for (int q = 0; q < PP_NQ; q++)
      v[q] = ((parts[(0 * PP_NQ + q) * N + p] + parts[(1 * PP_NQ + q) * N + p]) +
              parts[(2 * PP_NQ + q) * N + p]) + parts[(3 * PP_NQ + q) * N + p];
    gpx = v[0];
    gpy = v[1];
    Gs[2 * p] = v[2];
    Gs[2 * p + 1] = v[3];
    wsw[L.off_q + 2 * p] = xs[2 * p] + v[4];
    wsw[L.off_q + 2 * p + 1] = xs[2 * p + 1] + v[5];
    acc_ltf -= 0.5 * (v[2] * v[2] + v[3] * v[3]);
  }
  __syncthreads();

  // pass 2: three-body contraction (dotg, reptation.f:2928-2979), written per
  // particle i as g_i += sum_k M_ik (G_i - G_k),
  // h_i += -|sum_k M_ik|^2 - sum_k |M_ik|^2 + sum_k a_ik u_ik.(G_i - G_k)
  // with M = xi I + (xi'/r) u u^T and a = xi'' + 3 xi'/r; M and a.(u dG) are
  // even in the pair, M dG is odd
  if (s.has_u3) {
    for (int k = tid; k < 4 * PP_NQ * N; k += QMC_BLOCK) parts[k] = 0.0;
    __syncthreads();
    double ia[PP_NQ] = {0, 0, 0, 0, 0, 0};  // g3x, g3y, e11, e12, e22 of particle i
    int icur = -1;
    for (int u = u0; u < u1; u++) {
      int ti, tj, d;
      rr_unit(u, two, t1, ti, tj, d);
      if (ti != icur) {
        if (icur >= 0 && l < pick2(s.np, icur)) pp_add(part, N, pick2(s.ipf, icur) + l, ia, 5);
#pragma unroll
        for (int q = 0; q < PP_NQ; q++) ia[q] = 0.0;
        icur = ti;
      }
      const TpInfo* T = s.tpi + tpair(ti, tj);
      const TabRef tu3{T->row[TK_U3], T->ng[TK_U3]};
      const int jl = (l + d) & 63;
      if (tu3.row0 >= 0 && l < pick2(s.np, ti) && jl < pick2(s.np, tj) && (ti != tj || d < 32 || l < 32)) {
        const int i = pick2(s.ipf, ti) + l, j = pick2(s.ipf, tj) + jl;
        const double2 xi2 = ld2(xs, i), xj2 = ld2(xs, j);
        PairGeo p = pair_geo(s, xi2.x, xi2.y, xj2.x, xj2.y);
        double t, dt, ddt;
        spline3m<MIX>(s, tu3, p.ind, p.rem, t, dt, ddt);
        const double c = dt * p.byr;
        const double m11 = t + c * p.ux * p.ux, m12 = c * p.ux * p.uy, m22 = t + c * p.uy * p.uy;
        const double2 Gi = ld2(Gs, i), Gj = ld2(Gs, j);
        const double dGx = Gi.x - Gj.x, dGy = Gi.y - Gj.y;
        const double g3x = m11 * dGx + m12 * dGy, g3y = m12 * dGx + m22 * dGy;
        const double nm2 = m11 * m11 + 2.0 * m12 * m12 + m22 * m22;
        const double h3 = (ddt + 3.0 * c) * (p.ux * dGx + p.uy * dGy) - nm2;
        ia[0] += g3x;
        ia[1] += g3y;
        ia[2] += m11;
        ia[3] += m12;
        ia[4] += m22;
        const double ja[PP_NQ] = {-g3x, -g3y, m11, m12, m22, 0.0};
        pp_add(part, N, j, ja, 5);
        if (ti == 0) acc_h0 += h3; else acc_h1 += h3;
        if (tj == 0) acc_h0 += h3; else acc_h1 += h3;
      }
    }
    if (icur >= 0 && l < pick2(s.np, icur)) pp_add(part, N, pick2(s.ipf, icur) + l, ia, 5);
    __syncthreads();
    for (int p = tid; p < N; p += QMC_BLOCK) {
      double v[5];
#pragma unroll
      for (int q = 0; q < 5; q++)
        v[q] = ((parts[(0 * PP_NQ + q) * N + p] + parts[(1 * PP_NQ + q) * N + p]) +
                parts[(2 * PP_NQ + q) * N + p]) + parts[(3 * PP_NQ + q) * N + p];
      gpx += v[0];
      gpy += v[1];
      const double e2 = v[2] * v[2] + 2.0 * v[3] * v[3] + v[4] * v[4];
      if (ptype(s, p) == 0) acc_h0 -= e2; else acc_h1 -= e2;
    }
  }
  for (int p = tid; p < N; p += QMC_BLOCK) {
    wsw[L.off_gj + 2 * p] = gpx;
    wsw[L.off_gj + 2 * p + 1] = gpy;
  }

  // rho_k per type and the k-space Ewald sum
  double pk0 = 0.0, pk1 = 0.0, pkt = 0.0;
  if (s.nrhok) {
    double* pw = props + (int64_t)w * s.n_props;
    // passes of KMAXPT slots per thread: slots q = 2h, 2h+1 of a thread are
    // the two k-vectors of one item, which share their ky component
    for (int kb = 0; kb < s.nkslot; kb += KMAXPT * QMC_BLOCK) {
      double rc0[KMAXPT], rc1[KMAXPT], rs0[KMAXPT], rs1[KMAXPT];
      int kq[KMAXPT];              // k-vector of slot q (-1: none)
      int ox[KMAXPT], oy[KMAXPT];  // kx and ky phase slots of slot q (0 when none)
#pragma unroll
      for (int q = 0; q < KMAXPT; q++) {
        rc0[q] = rc1[q] = rs0[q] = rs1[q] = 0.0;
        const int n = kb + 2 * (threadIdx.x + (q >> 1) * QMC_BLOCK) + (q & 1);
        const int code = (n < s.nkslot && s.nkxv > 0) ? s.kxy[n] : 0;
        kq[q] = n >= s.nkslot ? -1 : (s.nkxv > 0 ? (code >> 16) - 1 : n);
        ox[q] = code & 0xff;
        oy[q] = ((code >> 8) & 0xff) + s.nkxv;
      }
      if (s.nkxv > 0) {
        // e^{i k.x} = e^{i kx x} e^{i ky y}: per particle chunk, the distinct
        // x and y phases go to LDS, then each thread multiplies its k's pairs
        using ph_t = typename std::conditional<MIX, float2, double2>::type;
        ph_t* ph = (ph_t*)(sm + 4 * N + 8);
        const int nv = s.nkxv + s.nkyv;
        for (int c0 = 0; c0 < N; c0 += s.kchunk) {
          const int cn = min(s.kchunk, N - c0);
          __syncthreads();
          if (s.kprog) {
            // runs of KRUN phases: the first directly, the rest by products
            // with e^{i dk x} (at most KRUN - 1 roundings from a direct value)
            const int gx = (s.nkxv + KRUN - 1) / KRUN, ng = gx + (s.nkyv + KRUN - 1) / KRUN;
            for (int e = threadIdx.x; e < cn * ng; e += QMC_BLOCK) {
              const int pc = e / ng, g = e - pc * ng;
              const int i = c0 + pc;
              const bool isx = g < gx;
              const int v0 = (isx ? g : g - gx) * KRUN, nrun = min(KRUN, (isx ? s.nkxv : s.nkyv) - v0);
              const double u = isx ? xs[2 * i] : xs[2 * i + 1];
              double sa, ca, sd, cd;
              qsincos_sreg((isx ? s.kxv[v0] : s.kyv[v0]) * u, &sa, &ca);
              qsincos_sreg((isx ? s.dkx : s.dky) * u, &sd, &cd);
              ph_t* dst = ph + pc * nv + (isx ? 0 : s.nkxv) + v0;
              dst[0] = ph_t{(decltype(ph_t::x))ca, (decltype(ph_t::x))sa};
#pragma unroll
              for (int r = 1; r < KRUN; r++) {
                if (r < nrun) {
                  const double c2 = fma(ca, cd, -sa * sd);
                  sa = fma(ca, sd, sa * cd);
                  ca = c2;
                  dst[r] = ph_t{(decltype(ph_t::x))ca, (decltype(ph_t::x))sa};
                }
              }
            }
          } else {
            for (int e = threadIdx.x; e < cn * nv; e += QMC_BLOCK) {
              const int pc = e / nv, v = e - pc * nv;
              const int i = c0 + pc;
              const double arg = v < s.nkxv ? s.kxv[v] * xs[2 * i] : s.kyv[v - s.nkxv] * xs[2 * i + 1];
              double sv, cv;
              qsincos_sreg(arg, &sv, &cv);
              ph[e] = ph_t{(decltype(ph_t::x))cv, (decltype(ph_t::x))sv};
            }
          }
          __syncthreads();
          // particles of type 0 come first: [0, split) of the chunk, each
          // type a loop of its own
          const int split = s.ntypes > 1 ? min(max(s.ipf[1] - c0, 0), cn) : cn;
          auto accum = [&](int pa, int pb, double (&rc)[KMAXPT], double (&rs)[KMAXPT]) {
            if constexpr (!MIX) {
              for (int pc = pa; pc < pb; pc++) {
                const ph_t* row = ph + pc * nv;
#pragma unroll
                for (int h = 0; h < KMAXPT; h += 2) {
                  // one ky phase for the item's two kx phases
                  const ph_t b = row[oy[h]], a0 = row[ox[h]], a1 = row[ox[h + 1]];
                  rc[h] = fma(a0.x, b.x, fma(-a0.y, b.y, rc[h]));
                  rs[h] = fma(a0.x, b.y, fma(a0.y, b.x, rs[h]));
                  rc[h + 1] = fma(a1.x, b.x, fma(-a1.y, b.y, rc[h + 1]));
                  rs[h + 1] = fma(a1.x, b.y, fma(a1.y, b.x, rs[h + 1]));
                }
              }
            } else {
              // the chunk's products summed in FP32, the chunk sum added to
              // the FP64 accumulators
              float qc[KMAXPT], qs[KMAXPT];
#pragma unroll
              for (int h = 0; h < KMAXPT; h++) qc[h] = qs[h] = 0.0f;
              for (int pc = pa; pc < pb; pc++) {
                const ph_t* row = ph + pc * nv;
#pragma unroll
                for (int h = 0; h < KMAXPT; h += 2) {
                  const ph_t b = row[oy[h]], a0 = row[ox[h]], a1 = row[ox[h + 1]];
                  qc[h] = fmaf(a0.x, b.x, fmaf(-a0.y, b.y, qc[h]));
                  qs[h] = fmaf(a0.x, b.y, fmaf(a0.y, b.x, qs[h]));
                  qc[h + 1] = fmaf(a1.x, b.x, fmaf(-a1.y, b.y, qc[h + 1]));
                  qs[h + 1] = fmaf(a1.x, b.y, fmaf(a1.y, b.x, qs[h + 1]));
                }
              }
#pragma unroll
              for (int h = 0; h < KMAXPT; h++) {
                rc[h] += (double)qc[h];
                rs[h] += (double)qs[h];
              }
            }
          };
          accum(0, split, rc0, rs0);
          accum(split, cn, rc1, rs1);
        }
      } else {
#pragma unroll
        for (int q = 0; q < KMAXPT; q++) {
          const int k = kq[q];
          if (k < 0) continue;
          const double kx = s.kvec[2 * k], ky = s.kvec[2 * k + 1];
          const int split = s.ntypes > 1 ? s.ipf[1] : N;
          for (int i = 0; i < split; i++) {
            double sv, cv;
            qsincos_sreg(kx * xs[2 * i] + ky * xs[2 * i + 1], &sv, &cv);
            rc0[q] += cv;
            rs0[q] += sv;
          }
          for (int i = split; i < N; i++) {
            double sv, cv;
            qsincos_sreg(kx * xs[2 * i] + ky * xs[2 * i + 1], &sv, &cv);
            rc1[q] += cv;
            rs1[q] += sv;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < KMAXPT; q++) {
        const int k = kq[q];
        if (k < 0) continue;
        if (!s.irhok[0]) rc0[q] = rs0[q] = 0.0;
        if (s.ntypes < 2 || !s.irhok[1]) rc1[q] = rs1[q] = 0.0;
        if (s.irhok[0]) {
          const int slot = s.jrhok + (s.irhok[0] - 1) * 2 * s.nk + 2 * k;
          pw[slot] = rc0[q];
          pw[slot + 1] = rs0[q];
        }
        if (s.ntypes > 1 && s.irhok[1]) {
          const int slot = s.jrhok + (s.irhok[1] - 1) * 2 * s.nk + 2 * k;
          pw[slot] = rc1[q];
          pw[slot + 1] = rs1[q];
        }
        // ordered type pairs (0,0), (0,1), (1,0), (1,1) (reptation.f:4034-4052)
        if (k < s.nke[0]) {
          const double aux = s.vk[0][k] * (rc0[q] * rc0[q] + rs0[q] * rs0[q]);
          pk0 += aux;
          pkt += aux;
        }
        if (s.ntypes > 1) {
          if (k < s.nke[1]) {
            const double aux = s.vk[1][k] * (rc0[q] * rc1[q] + rs0[q] * rs1[q]);
            pk0 += aux;
            pk1 += aux;
            pkt += 2.0 * aux;
          }
          if (k < s.nke[2]) {
            const double aux = s.vk[2][k] * (rc1[q] * rc1[q] + rs1[q] * rs1[q]);
            pk1 += aux;
            pkt += aux;
          }
        }
      }
    }
  }

  acc_ltf = block_sum(acc_ltf, red);
  acc_h0 = block_sum(acc_h0, red);
  acc_h1 = block_sum(acc_h1, red);
  acc_v0 = block_sum(acc_v0, red);
  acc_v1 = block_sum(acc_v1, red);
  pk0 = block_sum(pk0, red);
  pk1 = block_sum(pk1, red);
  pkt = block_sum(pkt, red);
  if (threadIdx.x == 0) {
    double c0 = 0.0, c1 = 0.0, ct = 0.0;
    for (int it = 0; it < s.ntypes; it++)
      for (int jt = 0; jt < s.ntypes; jt++) {
        const int tp = tpair(it, jt);
        if (pick3(s.nke, tp) != 0) {
          const double a = pick3(s.vk0, tp) * s.npnorm;
          if (it == 0) c0 += a;
          else c1 += a;
          ct += a;
        }
      }
    double* sc = wsw + L.off_sc;
    sc[SC_LTF] = acc_ltf;
    sc[SC_H0] = acc_h0;
    sc[SC_H1] = acc_h1;
    // the pair-potential tail constants are added in k_final (held here
    // through the pair loops, they had taken scalar registers the loops need)
    sc[SC_P0] = acc_v0 + pk0 + c0;
    sc[SC_P1] = acc_v1 + pk1 + c1;
    sc[SC_PT] = acc_v0 + acc_v1 + pkt + ct;
  }
}

__global__ __launch_bounds__(QMC_BLOCK, 5) void k_pairs(DevSys s, WsLayout L,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ ws,
                                                     double* __restrict__ props) {
  k_pairs_body<false>(s, L, x, ws, props);
}

__global__ __launch_bounds__(QMC_BLOCK, 5) void k_pairs_mix(DevSys s, WsLayout L,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ ws,
                                                         double* __restrict__ props) {
  k_pairs_body<true>(s, L, x, ws, props);
}

// ---------------------------------------------------------------------------
// k_det
// LDS: A[m*m] complex (column-major, row = orbital, column = particle; the
// inverse overwrites it with row = particle, column = orbital), then a union
// region reused by the phases.
__global__ __launch_bounds__(QMC_BLOCK) void k_det(DevSys s, WsLayout L,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ ws,
                                                   double* __restrict__ Fbuf,
                                                   double* __restrict__ Abuf, int kc,
                                                   int tpp_log2) {
  extern __shared__ __attribute__((aligned(16))) double sm[];
  __shared__ int s_piv[QMC_MAX_ORB];
  __shared__ int s_l, s_sing;
  __shared__ double s_pmax;
  __shared__ zd s_pv, s_akk;
  __shared__ double red[QMC_BLOCK / 64];
  const int w = blockIdx.x, sp = blockIdx.y;
  if (s.nact && w >= *s.nact) return;
  const int m = pick2(s.np, sp);
  if (m == 0) return;
  const int N = s.nptot, p0 = pick2(s.ipf, sp);
  const int tid = threadIdx.x;
  zd* A = (zd*)sm;
  double* U = sm + 2 * m * m;
  double* wsw = ws + (int64_t)w * L.per_walker;
  double* dsc = wsw + pick2(L.off_det, sp);
  zd* Ag = (zd*)(Abuf + (int64_t)w * L.a_per_walker + pick2(L.a_off, sp));  // row-major [k][j]
  zd* Fg = (zd*)(Fbuf + (int64_t)w * L.f_per_walker + pick2(L.f_off, sp));   // [beta][l][j]
  const double* qw = wsw + L.off_q;
  if (tid == 0) s_sing = 0;

  // 1. orbital matrix A(k,i) = exp(i g_k . q_i) (paho_orbitals, reptation.f:10298-10305)
  for (int e = tid; e < m * m; e += QMC_BLOCK) {
    const int k = e / m, i = e - k * m;
    const double kr = qw[2 * (p0 + i)] * s.gvec[2 * k] + qw[2 * (p0 + i) + 1] * s.gvec[2 * k + 1];
    double sv, cv;
    sincos(kr, &sv, &cv);
    const zd z = zmk(cv, sv);
    A[i * m + k] = z;
    Ag[k * m + i] = z;
  }
  __syncthreads();

  // 2. LU with partial pivoting, LINPACK zgefa semantics: pivot = first row
  // of maximal |re|+|im|; rows are swapped only right of the pivot column.
  for (int k = 0; k < m - 1; k++) {
    if (tid < 64) {
      const int r = k + tid;
      double v = -1.0;
      int idx = 0x7fffffff;
      if (r < m) { v = cabs1(A[k * m + r]); idx = r; }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(idx, o, 64);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
      }
      if (tid == 0) {
        s_l = idx;
        s_pmax = v;
        s_pv = A[k * m + idx];
        s_akk = A[k * m + k];
        s_piv[k] = idx;
        if (v == 0.0) s_sing = 1;
      }
    }
    __syncthreads();
    if (s_pmax == 0.0) {  // column already triangular (zgefa info = k)
      __syncthreads();
      continue;
    }
    const int l = s_l;
    const zd pv = s_pv, akk = s_akk;
    const zd tneg = zrcp(pv);
    const zd t = zmk(-tneg.x, -tneg.y);
    const int nr = m - k - 1;
    for (int e = tid; e < nr + nr + 1; e += QMC_BLOCK) {
      if (e < nr) {
        const int j = k + 1 + e;
        if (l != k) {
          const zd a = A[j * m + k];
          A[j * m + k] = A[j * m + l];
          A[j * m + l] = a;
        }
      } else {
        const int i = k + (e - nr);
        if (i == k) A[k * m + k] = pv;
        else {
          const zd old = (i == l) ? akk : A[k * m + i];
          A[k * m + i] = zmul(t, old);
        }
      }
    }
    __syncthreads();
    for (int e = tid; e < nr * nr; e += QMC_BLOCK) {
      const int jj = e / nr, ii = e - jj * nr;
      const int i = k + 1 + ii, j = k + 1 + jj;
      A[j * m + i] = zfma(A[j * m + k], A[k * m + i], A[j * m + i]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    s_piv[m - 1] = m - 1;
    if (cabs1(A[(m - 1) * m + (m - 1)]) == 0.0) s_sing = 1;
  }
  __syncthreads();
  // log|det| = sum log|U_kk| (zgedi det + log, reptation.f:10063, 9412-9430)
  {
    double v = 0.0;
    if (tid < m) v = log(hypot(A[tid * m + tid].x, A[tid * m + tid].y));
    v = block_sum(v, red);
    if (tid == 0) {
      dsc[DS_LOGDET] = v;
      dsc[DS_SING] = s_sing ? 1.0 : 0.0;
    }
  }
  if (s_sing) return;

  // 3. inverse (zgedi job=01 ordering, reptation.f:9437-9468)
  // 3a. inverse of U in place
  zd* rowk = (zd*)U;
  zd* colk = rowk + m;
  for (int k = 0; k < m; k++) {
    for (int e = tid; e < m; e += QMC_BLOCK) {
      if (e > k) rowk[e] = A[e * m + k];
      else if (e < k) colk[e] = A[k * m + e];
      else colk[e] = A[k * m + k];
    }
    __syncthreads();
    const zd tk = zrcp(colk[k]);
    const int ncol = m - k - 1;
    for (int e = tid; e < (k + 1) * (ncol + 1); e += QMC_BLOCK) {
      const int jj = e / (k + 1), i = e - jj * (k + 1);
      const zd ci = (i < k) ? zmul(zmk(-tk.x, -tk.y), colk[i]) : tk;
      if (jj == 0) {
        A[k * m + i] = ci;
      } else {
        const int j = k + jj;
        const zd base = (i < k) ? A[j * m + i] : zmk(0.0, 0.0);
        A[j * m + i] = zfma(rowk[j], ci, base);
      }
    }
    __syncthreads();
  }
  // 3b. inverse(U) * inverse(L), right to left, with the column swaps
  zd* work = (zd*)U;
  zd* part = work + m;  // [4][m]
  for (int k = m - 2; k >= 0; k--) {
    for (int i = tid; i < m; i += QMC_BLOCK)
      if (i > k) work[i] = A[k * m + i];
    __syncthreads();
    {
      const int g = tid >> 6, i = tid & 63;
      zd acc = zmk(0.0, 0.0);
      if (i < m)
        for (int j = k + 1 + g; j < m; j += 4) acc = zfma(work[j], A[j * m + i], acc);
      if (i < m) part[g * m + i] = acc;
    }
    __syncthreads();
    if (tid < m) {
      const int i = tid;
      zd v = (i <= k) ? A[k * m + i] : zmk(0.0, 0.0);
      for (int g = 0; g < 4; g++) v = zadd(v, part[g * m + i]);
      A[k * m + i] = v;
    }
    __syncthreads();
    const int l = s_piv[k];
    if (l != k) {
      for (int i = tid; i < m; i += QMC_BLOCK) {
        const zd a = A[k * m + i];
        A[k * m + i] = A[l * m + i];
        A[l * m + i] = a;
      }
    }
    __syncthreads();
  }

  // 4. F_b = A^-1 diag(i g^b) A (zv, reptation.f:10182-10194), 4x4 tiles;
  // A rows staged from the global copy kc at a time.
  {
    const int tl = tid & 15, tj = tid >> 4;
    const int l0 = 4 * tl, j0 = 4 * tj;
    zd Qx[4][4], Qy[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int c = 0; c < 4; c++) { Qx[r][c] = zmk(0, 0); Qy[r][c] = zmk(0, 0); }
    zd* Ac = (zd*)U;
    for (int k0 = 0; k0 < m; k0 += kc) {
      const int kn = min(kc, m - k0);
      for (int e = tid; e < kn * m; e += QMC_BLOCK) Ac[e] = Ag[k0 * m + e];
      __syncthreads();
      if (l0 < m && j0 < m) {
        for (int kk = 0; kk < kn; kk++) {
          const int k = k0 + kk;
          const double gx = s.gvec[2 * k], gy = s.gvec[2 * k + 1];
          zd ai[4], aj[4];
#pragma unroll
          for (int r = 0; r < 4; r++) ai[r] = (l0 + r < m) ? A[k * m + l0 + r] : zmk(0, 0);
#pragma unroll
          for (int c = 0; c < 4; c++) aj[c] = (j0 + c < m) ? Ac[kk * m + j0 + c] : zmk(0, 0);
#pragma unroll
          for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 4; c++) {
              const zd P = zmul(ai[r], aj[c]);
              Qx[r][c].x = fma(gx, P.x, Qx[r][c].x);
              Qx[r][c].y = fma(gx, P.y, Qx[r][c].y);
              Qy[r][c].x = fma(gy, P.x, Qy[r][c].x);
              Qy[r][c].y = fma(gy, P.y, Qy[r][c].y);
            }
        }
      }
      __syncthreads();
    }
    // F = i Q; store and keep the diagonal
    double* dl = U;  // d[j][b] complex: 4 doubles per j
    if (l0 < m && j0 < m) {
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int l = l0 + r, j = j0 + c;
          if (l < m && j < m) {
            const zd fx = zmk(-Qx[r][c].y, Qx[r][c].x);
            const zd fy = zmk(-Qy[r][c].y, Qy[r][c].x);
            Fg[(0 * m + l) * m + j] = fx;
            Fg[(1 * m + l) * m + j] = fy;
            if (l == j) {
              dl[4 * j + 0] = fx.x;
              dl[4 * j + 1] = fx.y;
              dl[4 * j + 2] = fy.x;
              dl[4 * j + 3] = fy.y;
            }
          }
        }
    }
  }
  // diag(A^-1 d2A) real parts: dd[j] = (xx, xy, yy)
  double* dd = U + 4 * m;
  double* xs = dd + 3 * m;
  for (int j = tid; j < m; j += QMC_BLOCK) {
    double sxx = 0, sxy = 0, syy = 0;
    for (int k = 0; k < m; k++) {
      const zd P = zmul(A[k * m + j], Ag[k * m + j]);
      const double gx = s.gvec[2 * k], gy = s.gvec[2 * k + 1];
      sxx -= gx * gx * P.x;
      sxy -= gx * gy * P.x;
      syy -= gy * gy * P.x;
    }
    dd[3 * j] = sxx;
    dd[3 * j + 1] = sxy;
    dd[3 * j + 2] = syy;
  }
  const double* xw = x + (int64_t)w * 2 * N;
  for (int k = tid; k < 2 * N; k += QMC_BLOCK) xs[k] = xw[k];
  __syncthreads();

  // 5. drift and the diagonal Laplacian terms for every particle
  // (reptation.f:10195-10240), with the backflow Jacobian qa and the
  // quasi-coordinate Laplacian qb regenerated from the pair table.
  const double* dl = U;
  const int tpp = 1 << tpp_log2;
  const int lane = tid & (tpp - 1), grp = tid >> tpp_log2, ngrp = QMC_BLOCK >> tpp_log2;
  double acc_g2 = 0.0, acc_h0 = 0.0, acc_h1 = 0.0;
  double* gdet = dsc + DS_N;
  double* Qd = gdet + 2 * N;
  for (int i0 = 0; i0 < N; i0 += ngrp) {
    const int ip = i0 + grp;
    const bool act = ip < N;
    const int ti = act ? ptype(s, ip) : 0;
    const bool on = act && (sp ? pick2(s.iub_flag[1], ti) : pick2(s.iub_flag[0], ti));
    const bool insp = ip >= p0 && ip < p0 + m;
    double zxr = 0, zxi = 0, zyr = 0, zyi = 0, hq = 0, hd = 0;
    double q11 = 0, q12 = 0, q22 = 0, bx = 0, by = 0;
    if (on) {
      const double xi = xs[2 * ip], yi = xs[2 * ip + 1];
      for (int c = lane; c < N; c += tpp) {
        if (c == ip) continue;
        const TabRef tb = tabref<TK_B>(s, tpair(ti, ptype(s, c)));
        if (tb.row0 < 0) continue;
        PairGeo p = pair_geo(s, xi, yi, xs[2 * c], xs[2 * c + 1]);
        double eta, deta, ddeta;
        spline3(s, tb, p.ind, p.rem, eta, deta, ddeta);
        const double cb = deta * p.byr;
        const double mb11 = eta + cb * p.ux * p.ux, mb12 = cb * p.ux * p.uy,
                     mb22 = eta + cb * p.uy * p.uy;
        const double cc = ddeta + 3.0 * cb;
        if (insp) {
          q11 += mb11; q12 += mb12; q22 += mb22;
          bx += cc * p.ux; by += cc * p.uy;
        }
        if (c >= p0 && c < p0 + m) {
          const int jf = c - p0;
          const double dxr = dl[4 * jf], dxi = dl[4 * jf + 1], dyr = dl[4 * jf + 2],
                       dyi = dl[4 * jf + 3];
          // qa(.,.,c,ip) = -mb
          zxr -= dxr * mb11 + dyr * mb12;
          zxi -= dxi * mb11 + dyi * mb12;
          zyr -= dxr * mb12 + dyr * mb22;
          zyi -= dxi * mb12 + dyi * mb22;
          // qb(.,ip,c) = -cc u
          hq -= cc * (dxr * p.ux + dyr * p.uy);
          const double a11 = mb11 * mb11 + mb12 * mb12, a12 = mb12 * (mb11 + mb22),
                       a22 = mb12 * mb12 + mb22 * mb22;
          hd += dd[3 * jf] * a11 + 2.0 * dd[3 * jf + 1] * a12 + dd[3 * jf + 2] * a22;
        }
      }
    }
    for (int o = 1; o < tpp; o <<= 1) {
      zxr += __shfl_xor(zxr, o, 64);
      zxi += __shfl_xor(zxi, o, 64);
      zyr += __shfl_xor(zyr, o, 64);
      zyi += __shfl_xor(zyi, o, 64);
      hq += __shfl_xor(hq, o, 64);
      hd += __shfl_xor(hd, o, 64);
      q11 += __shfl_xor(q11, o, 64);
      q12 += __shfl_xor(q12, o, 64);
      q22 += __shfl_xor(q22, o, 64);
      bx += __shfl_xor(bx, o, 64);
      by += __shfl_xor(by, o, 64);
    }
    if (act && lane == 0) {
      if (on) {
        if (insp) {
          const int jf = ip - p0;
          const double a11 = 1.0 + q11, a12 = q12, a22 = 1.0 + q22;
          const double dxr = dl[4 * jf], dxi = dl[4 * jf + 1], dyr = dl[4 * jf + 2],
                       dyi = dl[4 * jf + 3];
          zxr += dxr * a11 + dyr * a12;
          zxi += dxi * a11 + dyi * a12;
          zyr += dxr * a12 + dyr * a22;
          zyi += dxi * a12 + dyi * a22;
          hq += dxr * bx + dyr * by;
          const double b11 = a11 * a11 + a12 * a12, b12 = a12 * (a11 + a22),
                       b22 = a12 * a12 + a22 * a22;
          hd += dd[3 * jf] * b11 + 2.0 * dd[3 * jf + 1] * b12 + dd[3 * jf + 2] * b22;
          Qd[3 * jf] = q11;
          Qd[3 * jf + 1] = q12;
          Qd[3 * jf + 2] = q22;
        }
        gdet[2 * ip] = -zxr;
        gdet[2 * ip + 1] = -zyr;
        acc_g2 += pick2(s.hbs2m, ti) * (zxi * zxi + zyi * zyi);
        if (ti == 0) acc_h0 -= hq + hd;
        else acc_h1 -= hq + hd;
      } else {
        gdet[2 * ip] = 0.0;
        gdet[2 * ip + 1] = 0.0;
      }
    }
  }
  acc_g2 = block_sum(acc_g2, red);
  acc_h0 = block_sum(acc_h0, red);
  acc_h1 = block_sum(acc_h1, red);
  if (tid == 0) {
    dsc[DS_H0] = acc_h0;
    dsc[DS_H1] = acc_h1;
    dsc[DS_G2] = acc_g2;
  }
}

// ---------------------------------------------------------------------------
// k_final: one wavefront per walker.  delta is the current time step (the
// COMMON delta that kinetic's drift cutoff reads, reptation.f:2755).
__global__ __launch_bounds__(QMC_BLOCK) void k_final(DevSys s, WsLayout L, double* __restrict__ ws,
                                                     double* __restrict__ props,
                                                     double* __restrict__ gout,
                                                     double* __restrict__ hout,
                                                     double* __restrict__ grad2,
                                                     int32_t* __restrict__ status, int nw,
                                                     double delta) {
  const int w = blockIdx.x * (QMC_BLOCK / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nw || (s.nact && w >= *s.nact)) return;
  const int N = s.nptot;
  double* wsw = ws + (int64_t)w * L.per_walker;
  const double* sc = wsw + L.off_sc;
  double ltf = sc[SC_LTF];
  bool sing = false, guard = false;
  if (s.has_det) {
    for (int sp = 0; sp < s.ntypes; sp++) {
      if (pick2(s.np, sp) == 0) continue;
      const double* dsc = wsw + pick2(L.off_det, sp);
      ltf -= dsc[DS_LOGDET];
      if (dsc[DS_SING] != 0.0) sing = true;
    }
    // underflow guard of slater_excitations (reptation.f:10087-10092)
    guard = !(exp(-ltf * 2.0) > 0.0);
  }
  const bool use_det = s.has_det && !guard && !sing;
  double h[2] = {sc[SC_H0], sc[SC_H1]};
  double g2ph = 0.0;
  if (use_det) {
    for (int sp = 0; sp < s.ntypes; sp++) {
      if (pick2(s.np, sp) == 0) continue;
      const double* dsc = wsw + pick2(L.off_det, sp);
      h[0] += dsc[DS_H0] + dsc[DS_X0];
      h[1] += dsc[DS_H1] + dsc[DS_X1];
      g2ph += dsc[DS_G2];
    }
  }
  double* gw = gout + (int64_t)w * 2 * N;
  double g2t[2] = {0.0, 0.0}, ggs = 0.0;
  for (int i = lane; i < N; i += 64) {
    double gx = wsw[L.off_gj + 2 * i], gy = wsw[L.off_gj + 2 * i + 1];
    if (use_det) {
      for (int sp = 0; sp < s.ntypes; sp++) {
        if (pick2(s.np, sp) == 0) continue;
        const double* gd = wsw + pick2(L.off_det, sp) + DS_N;
        gx += gd[2 * i];
        gy += gd[2 * i + 1];
      }
    }
    const int t = ptype(s, i);
    const double x2 = gx * gx + gy * gy;
    if (s.adrift != 0.0) {
      // drift cutoff (kinetic, reptation.f:2747-2760), vari as sigma sets it
      // (reptation.f:855-870); every particle here has hbs2m != 0
      const double vari = 1.0 / (2.0 * pick2(s.hbs2m, t) * delta);
      double gg = x2 * s.adrift * delta * delta * vari;
      gg = (-1.0 + sqrt(1.0 + 2.0 * gg)) / gg;
      ggs += gg;
      gx *= gg;
      gy *= gg;
    }
    if (t == 0) g2t[0] += x2;
    else g2t[1] += x2;
    gw[2 * i] = gx;
    gw[2 * i + 1] = gy;
  }
  g2t[0] = wave_sum(g2t[0]);
  g2t[1] = wave_sum(g2t[1]);
  ggs = wave_sum(ggs);
  if (lane == 0) {
    double* p = props + (int64_t)w * s.n_props;
    double kin0 = 0.0;
#pragma unroll
    for (int t = 0; t < 2; t++) {
      if (t >= s.ntypes) break;
      hout[(int64_t)w * 2 + t] = h[t];
      if (s.hbs2m[t] == 0.0) continue;  // no kinetic slots (reptation.f:2739)
      const double kin = s.hbs2m[t] * (h[t] - g2t[t]);
      kin0 += kin;
      p[s.jkin[t]] = kin / s.np[t];
      p[s.jun[t]] = 1.0 / sqrt(g2t[t]);
    }
    if (s.ntypes == 1) hout[(int64_t)w * 2 + 1] = 0.0;
    p[s.jkin0] = kin0 / s.npnorm;
    p[s.jgg] = (s.adrift == 0.0) ? 1.0 : ggs / s.npnorm;
    if (use_det) grad2[w] = g2ph;
    const double g2v = grad2[w];
    p[s.jefp] = (s.ntheta != 0) ? g2v / s.npnorm : 0.0;
    p[s.jltf] = ltf;
#pragma unroll
    for (int t = 0; t < 2; t++)
      if (t < s.ntypes && s.jpot[t] >= 0) p[s.jpot[t]] = (s.tailc[t] + sc[SC_P0 + t]) / s.np[t];
    p[s.jpot0] = (s.tailc_tot + sc[SC_PT] + s.v0) / s.npnorm;
    p[s.jetot] = p[s.jpot0] + p[s.jkin0] + p[s.jefp];
    p[s.je2] = p[s.jetot] * p[s.jetot];
    if (status) status[w] = (guard ? 1 : 0) | (sing ? 2 : 0);
  }
}
