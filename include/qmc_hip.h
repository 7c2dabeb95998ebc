/*
 * qmc_hip.h -- C-ABI of the MI355X walker-evaluation library (lib2degas_hip.so).
 *
 * Plain C: scalars by value, arrays by pointer, int return codes (0 = ok),
 * no allocation visible to the caller, no torch or HIP types.  Every entry
 * point replaces one seam of the reference Fortran (fabioaffinito/2degas);
 * the citation in front of each declaration names it.  The Fortran side binds
 * these through ISO_C_BINDING (2degas_amd/fortran/qmc_hip_mod.f90).
 *
 * Units and layouts follow the reference: energies in Rydberg, lengths in
 * units of rs, positions x[walker][particle][2] with all particles of type 1
 * first (the reference's x_new(2,mnp) column-major == C [particle][2]).
 */
#ifndef QMC_HIP_H
#define QMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
enum {
  QMC_OK = 0,
  QMC_ERR_ARG = 1,         /* bad argument / unsupported deck feature       */
  QMC_ERR_LIMIT = 2,       /* size above a compiled limit                   */
  QMC_ERR_SINGULAR = 3,    /* orbital matrix singular (reptation.f:10318)   */
  QMC_ERR_HIP = 4,         /* HIP runtime error                             */
  QMC_ERR_NOTWIST = 5,     /* determinant requested but no g-vectors set    */
  QMC_ERR_STATE = 6        /* call order (e.g. sweep before init)           */
};

/* Compiled limits of the device path. */
#define QMC_MAX_PARTICLES 128  /* total particles per walker (2 types x 64)  */
#define QMC_MAX_ORBITALS 64    /* particles per type in one determinant      */
#define QMC_MAX_TYPES 2

/*
 * Deck tables, copied once into the device context.  Mirrors the reference
 * COMMON blocks filled by `input` (reptation.f:12-710):
 *   /c_sys/ np, hbs2m, ntypes (ewald.h:124-130), /c_box/ el (ewald.h:150-151),
 *   /c_tables/ ut, tail, iexp, iu2table, iu3table, iubtable, iv2table
 *   (ewald.h:101-116), /c_dist/ drt, ngrid (ewald.h:84-99),
 *   /c_kspace/ kvec, nk (ewald.h:160-164), /c_ewald/ nk_ewald (ewald.h:272),
 *   v0, adrift (ewald.h:175-179), ntheta (ewald.h:288-290),
 *   /c_rhok/ irhok, nrhok (ewald.h:72-76).
 * `ut` keeps the reference layout ut(0:ldut-1, 4, ntab), column-major:
 * coefficient c (0..3) of row i of table t (0-based) is ut[(t*4+c)*ldut+i].
 * Rows ngrid+1..ngrid+nk_ewald of column 0 hold vk and row ngrid+nk_ewald+1
 * holds vk0 for a table read with a `kspace` block (reptation.f:4716-4724).
 * Table selectors are 1-based table numbers (0 = none), indexed [it*2+jt].
 */
typedef struct qmc_tables {
  int32_t ndim;            /* must be 2                                    */
  int32_t ntypes;          /* 1 or 2                                       */
  int32_t np[2];           /* particles per type                           */
  double hbs2m[2];         /* hbar^2/2m per type (reference units)         */
  double el[2];            /* box sides                                    */
  double drt;              /* radial grid step shared by all tables        */
  int32_t ntab;            /* number of tables                             */
  int32_t ldut;            /* leading dimension of ut                      */
  const double* ut;        /* [ntab][4][ldut]                              */
  const int32_t* ngrid;    /* [ntab]                                       */
  const double* tail;      /* [ntab]                                       */
  const int32_t* iexp;     /* [ntab]                                       */
  const int32_t* nk_ewald; /* [ntab]                                       */
  int32_t iu2[4], iu3[4], iub[4], iv2[4];
  double v2value[4];       /* pair-potential multipliers (reference: 1.0)  */
  double v0;               /* constant potential (`v0` keyword)            */
  int32_t nk;              /* k-vectors read by `kspace`                   */
  const double* kvec;      /* [nk][2]                                      */
  int32_t irhok[2];        /* rhok slot of each type, 0 = not requested    */
  int32_t nrhok;
  int32_t ntheta;          /* twists per direction (0 = Gamma, no efp)     */
  double adrift;           /* drift cutoff (0 = off, reference default)    */
  double delta;            /* initial time step of the drift cutoff; later
                              values come from qmc_set_delta and the sweeps */
} qmc_tables;

/* Sizes of a context (slot numbers are 1-based like the reference's). */
typedef struct qmc_info {
  int32_t nptot;           /* total particles                              */
  int32_t npnorm;          /* moving particles                             */
  int32_t n_scalar_props;  /* scalar property slots (reptation.f:577-616)  */
  int32_t n_props;         /* scalars + 2*nk per rhok type                 */
  int32_t jetot, jltf, jacc, jpot0, jkin0, je2, jgg, jefp, jrhok;
  int32_t jpot[2], jkin[2], jun[2];
  int32_t ng;              /* g-vectors currently set                       */
  int32_t has_det;         /* backflow determinant active                  */
  int32_t reserved0;       /* zero (aligns the byte counts below)          */
  /* Device memory, for callers that budget it (admission of jobs): what
   * the context holds now (tables, g-vectors and every per-walker buffer
   * allocated so far, each kept at its largest size), and what each walker
   * adds to the buffers that grow with the batch: the evaluation workspace
   * (every entry point), the VMC proposal buffers (qmc_vmc_*, qmc_rmc_*) and
   * the staging of the host-pointer qmc_eval_batch.  The walker state of
   * qmc_walkers is the caller's and is not counted.                        */
  uint64_t dev_bytes;
  uint64_t eval_bytes_per_walker;
  uint64_t vmc_bytes_per_walker;
  uint64_t stage_bytes_per_walker;
} qmc_info;

typedef struct qmc_ctx qmc_ctx;

/* Precision of an evaluation. */
enum { QMC_PREC_FP64 = 0, QMC_PREC_MIXED = 1 };

/*
 * Create a device context from deck tables on HIP device `device`.
 * Replaces the state set up by `input` (reptation.f:12-710) that
 * compute_properties reads.  Validates the deck: every pair of types with
 * particles must carry the same table kinds (the reference's pair counters
 * assume it, reptation.f:2993-3038).
 */
int qmc_ctx_create(const qmc_tables* tables, int device, qmc_ctx** out);
int qmc_ctx_destroy(qmc_ctx* ctx);
int qmc_ctx_info(const qmc_ctx* ctx, qmc_info* out);

/*
 * Set the g-vectors of the determinant orbitals (first np(it) are occupied
 * per spin).  Replaces the gvec/gtens COMMON written by `input`
 * (reptation.f:674-708) or by `tabc`/`twistg` (reptation.f:9559-9638,
 * 9818-9861).  gvec is [ng][2].
 */
int qmc_set_twist(qmc_ctx* ctx, const double* gvec, int ng);

/*
 * Host helper: the twisted, |g|^2-sorted g-vector list exactly as
 * twistg(ith,jth) builds it, including the reference's heap sort SORTTF
 * (reptation.f:9514-9557, 9818-9861).  With ntheta == 0 it returns the
 * untwisted list of `input` (reptation.f:674-708).  gvec_out holds 2*nk+1
 * vectors.
 */
int qmc_twist_gvec(const double* kvec, int nk, const double* el, int ntheta,
                   int ith, int jth, double* gvec_out);

/*
 * Set the time step that the drift cutoff of `kinetic` reads when the deck
 * sets adrift != 0: the COMMON delta of /c_sim/ that each driver record
 * assigns before calling sigma (reptation.f:740-746, 795-805, 855-870,
 * 2755).  qmc_eval_batch[_device] and qmc_vmc_start use the value set last;
 * qmc_vmc_sweep and qmc_rmc_start/qmc_rmc_sweep set it from their own
 * delta.  A Fortran forwarder passes its COMMON delta before every call.
 */
int qmc_set_delta(qmc_ctx* ctx, double delta);

/*
 * Batched local-energy evaluation: for each of nw walkers, everything
 * compute_properties(1) (reptation.f:2661-2727) writes:
 *   props[w][n_props]  p_new(1..n_props) (slot jacc is left as passed in),
 *   g[w][nptot][2]     g_new  = grad(-log psi),
 *   h[w][2]            h_new  = per-type laplacian of -log psi,
 *   grad2_ph[w]        fixed-phase |Im grad log D|^2 sum (in/out: left as
 *                      passed when the determinant underflow guard of
 *                      reptation.f:10087-10092 fires, as the reference does).
 * Host pointers; the library stages them through device memory.
 * status[w] (optional, may be NULL) gets 1 where the underflow guard fired.
 */
int qmc_eval_batch(qmc_ctx* ctx, int nw, const double* x, double* props,
                   double* g, double* h, double* grad2_ph, int32_t* status);

/*
 * Same as qmc_eval_batch with all arrays already in device memory and the
 * work queued on `stream` (a hipStream_t, NULL = default stream); returns
 * once queued.  `precision` is QMC_PREC_FP64, or QMC_PREC_MIXED: the u2, u3
 * and pair-potential tables evaluated in FP32 and the rho_k phases formed
 * and summed in FP32 over chunks of particles, each chunk's sum added into
 * FP64; the geometry, backflow, determinant and every accumulator stay FP64
 * (local energy within about 1e-6 relative of FP64 on sampled
 * configurations).
 */
int qmc_eval_batch_device(qmc_ctx* ctx, int nw, const double* d_x,
                          double* d_props, double* d_g, double* d_h,
                          double* d_grad2_ph, int32_t* d_status, int precision,
                          void* stream);

/*
 * qmc_eval_batch_device without any allocation: no counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * evaluation's workspace lives in a file-scope device array of
 * QMC_NOALLOC_SCRATCH_BYTES per GPU instead of context-owned memory; the
 * caller passes inputs and outputs as for qmc_eval_batch_device.  The array
 * is resolved once per device when a context is created on it, holds what
 * the last call left, and is shared by every context on that GPU: calls
 * that overlap on one GPU (several streams or threads) race.  Returns
 * QMC_ERR_LIMIT when nw walkers of this context need more scratch than the
 * limit (at N = 122, 256 KiB per walker: about 260 walkers), and does no
 * per-kernel timing.
 */
#define QMC_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define QMC_NOALLOC_MAX_DEVICES 64
int qmc_eval_batch_device_noalloc(qmc_ctx* ctx, int nw, const double* d_x,
                                  double* d_props, double* d_g, double* d_h,
                                  double* d_grad2_ph, int32_t* d_status, int precision,
                                  void* stream);

/*
 * Walker state of a batched VMC run, all in device memory and owned by the
 * caller (nw walkers; N = nptot).  Walker w plays the role of MPI rank w of
 * the reference: its own old configuration (/c_old/, ewald.h:49-52), age,
 * fixed-phase scratch and its own stream of the 48-bit LCG rannyu
 * (reptation.f:5247-5291), stored as the 48-bit integer
 * l1*2^36 + l2*2^24 + l3*2^12 + l4 of the four limbs.
 */
typedef struct qmc_walkers {
  int32_t nw;
  double* x;        /* [nw][N][2]  x_old                                   */
  double* g;        /* [nw][N][2]  g_old                                   */
  double* h;        /* [nw][2]     h_old                                   */
  double* props;    /* [nw][n_props] p_old                                 */
  double* grad2_ph; /* [nw]        fixed-phase term of the last evaluation */
  double* age;      /* [nw]        consecutive rejections                  */
  int64_t* rng;     /* [nw]        LCG state                               */
  double* blk;      /* [nw][n_props] block sums of p_old (averages(2)), or NULL */
} qmc_walkers;

/*
 * Start a VMC run from the configurations already in w->x: evaluate them and
 * accept unconditionally, as read_conf does (reptation.f:2486-2495:
 * compute_properties(1), then metropolis_test(1)).  Device pointers, queued
 * on `stream`.
 */
int qmc_vmc_start(qmc_ctx* ctx, const qmc_walkers* w, void* stream);

/*
 * nsteps VMC steps for every walker: move (reptation.f:4617-4680: Box-Muller
 * drift-diffusion proposal of all electrons, evaluation, reverse-move weight)
 * then metropolis_test (reptation.f:2622-2659: a uniform is drawn only when
 * p < 1; accept if p >= u; copy new into old), then averages(2)
 * (reptation.f:4300-4305: blk += p_old) when w->blk is not NULL.
 * delta is the time step (sigma, reptation.f:855-870: var = 2 hbs2m delta);
 * mage forces acceptance after that many consecutive rejections.
 * bm_order selects which rannyu() of the Box-Muller expression is drawn
 * first (0: the cos() argument, as the amdflang build of the reference does).
 * trace_p / trace_age (optional, [nsteps][nw]) receive p and age per step.
 */
int qmc_vmc_sweep(qmc_ctx* ctx, const qmc_walkers* w, int nsteps, double delta,
                  double mage, int bm_order, double* trace_p, double* trace_age,
                  void* stream);

/*
 * Wait for the queued VMC work and copy each walker's sticky status (bit 1:
 * underflow guard fired at some step, bit 2: singular orbital matrix) to
 * host memory; returns QMC_ERR_SINGULAR if any walker hit a singular matrix
 * (where the reference stops, reptation.f:10318).
 */
int qmc_vmc_status(qmc_ctx* ctx, int nw, int32_t* sticky);

/*
 * Reptation QMC (the reference's rmc driver, reptation.f:1286-1313).  A
 * reptile is a path of ntau beads held in a ring of nbuf = ntau + mdelta
 * slots (the reference's x_stack/g_stack/p_stack/age_stack/w_stack with
 * n_buffer slots, read_path at reptation.f:1766-1802).  Ring slots and the
 * path ends are 0-based here (the reference's jfirst - 1, jlast - 1).  The
 * bead sign s_stack is not stored: s_new is 1 on this path
 * (reptation.f:3057), so the fixed-node rejection of reptation.f:1985 never
 * fires.  All pointers are device pointers.
 */
typedef struct qmc_reptiles {
  int32_t nr;       /* reptiles                                              */
  int32_t ntau;     /* beads per path                                        */
  int32_t nbuf;     /* ring slots, ntau + mdelta                             */
  double* x;        /* [nr][nbuf][N][2]      bead positions (x_stack)        */
  double* g;        /* [nr][nbuf][N][2]      bead drifts (g_stack)           */
  double* props;    /* [nr][nbuf][n_props]   bead properties (p_stack)       */
  double* age;      /* [nr][nbuf]            bead ages (age_stack)           */
  double* w;        /* [nr][nbuf][5]         link weights (w_stack: jbra,
                       jdtp, jitp, jlng, jnds)                               */
  int32_t* ends;    /* [nr][2]               jfirst, jlast                   */
  double* rage;     /* [nr]                  reptile age (age)               */
  double* pp_old;   /* [nr][n_props]         path averages (p_p_old)         */
  double* grad2_ph; /* [nr]                  fixed-phase term, carried from
                       one evaluation to the next as the COMMON value is     */
  int64_t* rng;     /* [nr]                  LCG state                       */
  double* blk;      /* [nr][n_props]         block sums of p_old, or NULL    */
} qmc_reptiles;

/* Driver settings of the rmc record and the deck keywords that shape the
 * move (reptation.f:795-805, 46-53, 149-170). */
typedef struct qmc_rmc_params {
  double delta;      /* time step (var = 2 hbs2m delta)                      */
  double etrial;     /* trial energy in the branching weight                 */
  double mage;       /* bead age that forces a Langevin acceptance           */
  double mage_r;     /* reptile age that forces a reptation acceptance       */
  int32_t ntauskip;  /* beads skipped at each end by path_props              */
  int32_t mdelta;    /* largest number of beads moved at once                */
  int32_t rejection; /* 0: Langevin steps are always accepted                */
  int32_t wpiu;      /* propagator options of reptation_move                 */
  int32_t gstorto;
  int32_t nodalaction;
  int32_t ecut;      /* fgg: e when 0, value_ecut + (e - value_ecut) gg      */
  int32_t bm_order;  /* Box-Muller draw order, as for qmc_vmc_sweep         */
  double value_ecut;
} qmc_rmc_params;

/*
 * Start reptiles from the beads already in slots 0..ntau-1 of r->x (read_path,
 * reptation.f:1766-1802): each bead is evaluated in order and accepted as
 * read_conf does (age 0, p(jacc) = 1), the links get their weights and
 * Langevin probabilities (get_w_stack, reptation.f:1804-1851), the path
 * averages are formed (path_props, reptation.f:1556-1577) and the ends set to
 * 0 and ntau-1.
 */
int qmc_rmc_start(qmc_ctx* ctx, const qmc_reptiles* r, const qmc_rmc_params* prm, void* stream);

/*
 * nsteps reptation moves for every reptile: reptation_move
 * (reptation.f:1897-2113: direction and number of beads drawn, each bead a
 * drift-diffusion proposal from the path end, evaluated and accepted by the
 * Langevin test, then the path weight of the move) and
 * reptation_metropolis_test (reptation.f:1865-1895), then averages(2)
 * (blk += p_old) when r->blk is not NULL.  The random numbers of reptile k
 * are those of MPI rank k of the reference.  trace (optional, device,
 * [nsteps][nr][8 + mdelta]) receives per move: idir, ndelta, path weight,
 * reptile age after the test, the path averages of the local energy,
 * -log psi and the kinetic energy, spare, then the Langevin probability of
 * each bead moved (-1 past ndelta).
 */
int qmc_rmc_sweep(qmc_ctx* ctx, const qmc_reptiles* r, const qmc_rmc_params* prm, int nsteps,
                  double* trace, void* stream);

/*
 * Diffusion Monte Carlo with fixed-population branching (the reference's dmc
 * driver, reptation.f:872-921, and branch2, reptation.f:2162-2271), each
 * walker playing one MPI rank that holds one configuration (ntarget 1): its
 * moves draw its rank's stream, the comb of branch2 draws the second
 * generator (rannyu2, seeded alike on every rank, reptation.f:36-37), and a
 * walker whose multiplicity comes out 0 takes a copy of a walker with extra
 * copies, as the reference's send/receive between ranks does.
 *
 * qmc_dmc_params: the dmc record (reptation.f:768-777) for one call.
 * qmc_dmc_state: device buffers owned by the caller, for nw walkers.
 */
typedef struct qmc_dmc_params {
  double delta;      /* time step                                          */
  double etrial;     /* trial energy of the branching weights               */
  double mage;       /* consecutive rejections that force a move            */
  int32_t mmstep;    /* moves per walker between branchings                 */
  int32_t mmult;     /* bound of one move's weight factor (peso)            */
  int32_t bm_order;  /* Box-Muller draw order, as for qmc_vmc_sweep         */
  int32_t reserved0; /* zero                                                */
} qmc_dmc_params;

typedef struct qmc_dmc_state {
  double* wate;     /* [nw] weight of the current step (branch2's wate)     */
  double* wsum;     /* [nw] sums of the weights of averages(2) (blk_norm)   */
  int32_t* src;     /* [nw] walker each one was copied from by the last
                       branching (itself when kept), 0-based              */
  int64_t* rng2;    /* [1]  state of the second generator                   */
  int32_t* counts;  /* [4]  added to: weight factors bounded at mmult (all
                       walkers), weights bounded at 10 (all walkers; branch2's
                       nmult), of them walker 0's (rank 0's nmult), moves of
                       walker 0 forced by mage (rank 0's nage)             */
} qmc_dmc_state;

/*
 * nsteps DMC steps for every walker: mmstep times move + metropolis_test
 * (as qmc_vmc_sweep), the local energy of the step (the mean of the old and
 * new energies when accepted, reptation.f:2631), the weight factor
 * peso = exp(-npnorm (elocal - etrial) delta) bounded at mmult, wate *= peso
 * and averages(2) with weight wate (w->blk += p_old wate, d->wsum += wate;
 * w->blk must be set); then branch2 as qmc_dmc_branch and the copies.
 * d->wate is set to 1 at the start of every step.
 */
int qmc_dmc_sweep(qmc_ctx* ctx, const qmc_walkers* w, const qmc_dmc_state* d, int nsteps,
                  const qmc_dmc_params* prm, void* stream);

/*
 * branch2 alone on nw weights (device): weights above 10 are set to 10
 * (counts[1], and counts[2] for weight 0), the comb of nint(sum of the
 * fractional parts) rannyu2 draws on the normalized weights gives each
 * walker its multiplicity, and the walkers with multiplicity 0, in
 * ascending order, are given the extra copies of the walkers with more
 * than one, in ascending order: src[j] is where walker j's next
 * configuration comes from.  d_wate is overwritten.
 */
int qmc_dmc_branch(qmc_ctx* ctx, int nw, double* d_wate, int64_t* d_rng2, int32_t* d_src,
                   int32_t* d_counts, void* stream);

/*
 * Single-electron VMC moves with Sherman-Morrison updates of the inverse
 * Slater matrices.  No counterpart in the reference, whose move displaces
 * every electron at once (move, reptation.f:4617-4680).  For trial
 * functions without backflow (every backflow table zero) moving one
 * electron changes one column of its spin's matrix A(k, j) = exp(i g_k . r_j)
 * and the inverse is updated (the two- and three-body Jastrow factors are
 * both supported).  With backflow every quasi-coordinate moves with the
 * electron, no low-rank update applies, and each move evaluates the
 * proposed configuration's log|psi| (Jastrow sums, quasi-coordinates and
 * the determinants' pivots): there ainv is not used (may be NULL),
 * qmc_sem_refresh does nothing, and the walkers' ltf slot of props must
 * hold the current value (qmc_vmc_start) and is kept current.
 *
 * qmc_sem_state: device buffers owned by the caller, for w->nw walkers.
 * ainv holds per walker and spin the inverse, row = particle, column =
 * orbital, mm x mm with mm the larger spin's particle count: complex
 * double (re, im) with QMC_PREC_FP64, complex float with QMC_PREC_MIXED.
 */
typedef struct qmc_sem_state {
  void* ainv;       /* [nw][2][mm][mm] complex                              */
  double* dlogpsi;  /* [nw] added to: log|psi'| - log|psi| of each accepted
                       move (log|R| - dU)                                   */
  double* nacc;     /* [nw] added to: accepted moves                        */
} qmc_sem_state;

/* The inverses from the current positions w->x (FP64 Gauss-Jordan with
 * partial pivoting), stored at the precision asked for. */
int qmc_sem_refresh(qmc_ctx* ctx, const qmc_walkers* w, const qmc_sem_state* d, int precision,
                    void* stream);

/*
 * Moves of electrons j0..j1-1 (0-based), one after the other, for every
 * walker: a Gaussian step of width sigma per dimension (Box-Muller from the
 * walker's rannyu stream w->rng: two draws for the step, a third for the
 * acceptance), accepted with probability |R|^2 exp(-2 dU), where R is the
 * determinant ratio (A^-1 a')_j and dU the change of the Jastrow exponent
 * (the u2 sum and the three-body term of the full evaluation); with
 * backflow, with probability exp(-2 (ltf' - ltf)) from the proposal's
 * value.  Positions and inverses are updated in place; w->g, w->h and the
 * other slots of w->props are stale until the next evaluation.
 */
int qmc_sem_moves(qmc_ctx* ctx, const qmc_walkers* w, const qmc_sem_state* d, int j0, int j1,
                  double sigma, int precision, void* stream);

/*
 * qmc_sem_moves for up to 16 jobs (contexts of any N, decks and walker sets
 * on one device, none with backflow) in one launch per spin: spin 0's moves
 * of every job, then spin 1's.  Each job's results are those of its own
 * qmc_sem_moves call with the same range, bit for bit.  No counterpart in
 * the reference (one system per run); for services that run many small
 * jobs at once.
 */
typedef struct qmc_sem_job {
  qmc_ctx* ctx;
  const qmc_walkers* w;
  const qmc_sem_state* d;
  int32_t j0, j1;    /* electrons j0..j1-1 of every walker of the job        */
  double sigma;      /* Gaussian step per dimension                           */
} qmc_sem_job;

int qmc_sem_moves_packed(const qmc_sem_job* jobs, int njobs, int precision, void* stream);

/*
 * Tile shape of the single-electron move kernel for a context (no
 * counterpart in the reference): threads per row of the inverse (4, or 2
 * for the 32- and 64-row classes; a row class holds spins of up to 16, 32
 * or 64 particles) at a precision.  Defaults: 2 for 32 rows, 4 otherwise.
 * The walk depends on it only through the order of the Jastrow sums
 * (rounding); an autotuner times the shapes (SingleMoves.tune_tile in the
 * Python binding) and sets the faster.
 */
int qmc_sem_set_tile(qmc_ctx* ctx, int precision, int rows, int threads_per_row);

/*
 * Per-kernel timing with HIP events recorded on the launch stream around each
 * kernel of every evaluation while enabled (no counterpart in the reference,
 * which has no timers).  qmc_timing(ctx, 1) resets and enables,
 * qmc_timing(ctx, 0) disables.  qmc_timing_read waits for the recorded work
 * and returns the mean milliseconds per launch of the kernels in the order
 * pairs, determinant (k_gjm, k_gjm_part or k_det), drift and diagonal terms
 * (k_fgrad; 0 when k_det does them), cross term, final (ms[5]) and the
 * launches counted.
 */
int qmc_timing(qmc_ctx* ctx, int enable);
int qmc_timing_read(qmc_ctx* ctx, double* ms, int* nlaunch);

/*
 * Autotuner of the pair kernels' phase chunk (no counterpart in the
 * reference): times full evaluations of nw synthetic configurations
 * (uniform in the box) on `stream` for the size-based default chunk and
 * chunks of 1, 2, 4, ... 32 particles (those whose LDS fits in 64 KB), for
 * the FP64 and the mixed-precision
 * pair kernel, and keeps the fastest for the context.  kchunk_out[2] and
 * ms_out[2] (optional) receive the chosen chunks and their ms per
 * evaluation.  Different chunks add the rho_k phases in a different order
 * (rounding only).  Decks whose k-vectors do not factor into distinct
 * components have nothing to tune (kchunk_out 1).
 */
int qmc_autotune(qmc_ctx* ctx, int nw, int32_t* kchunk_out, double* ms_out, void* stream);

/* Human-readable text of an error code, and of the last HIP error seen. */
const char* qmc_strerror(int code);
const char* qmc_last_error(void);

/*
 * Algorithmic floating-point work of one local-energy evaluation for a
 * system of nup+ndown particles with nk_ewald k-vectors (real flops, FMA=2):
 * the per-evaluation count used for roofline fractions.
 */
double qmc_flops_per_eval(int nup, int ndown, int nk_ewald);

#ifdef __cplusplus
}
#endif
#endif /* QMC_HIP_H */
