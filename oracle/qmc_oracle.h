/*
 * qmc_oracle.h -- CPU FP64 restatement of the reference's walker-evaluation
 * path.  TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg as the checker.  The product library never
 * links or calls it.
 *
 * Every function restates a reference routine (file:line given in
 * qmc_oracle.c) in plain C with the reference's loop order, so it pins the
 * HIP path to the reference's arithmetic.  It shares only the qmc_tables /
 * qmc_info structs of include/qmc_hip.h.
 */
#ifndef QMC_ORACLE_H
#define QMC_ORACLE_H
#include <stdint.h>
#include "../include/qmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc* orc_create(const qmc_tables* t, int* err);
void orc_destroy(orc* o);
int orc_info(const orc* o, qmc_info* out);
int orc_set_twist(orc* o, const double* gvec, int ng);
/* variant 0: tabc's theta association (reptation.f:9578-9583);
 * variant 1: twistg's (reptation.f:9823-9824). */
int orc_twist_gvec(const double* kvec, int nk, const double* el, int ntheta,
                   int ith, int jth, double* gvec_out, int variant);

/* Current time step read by the drift cutoff of kinetic when adrift != 0
 * (the COMMON delta, reptation.f:2755). */
int orc_set_delta(orc* o, double delta);

/* compute_properties(1) for one walker (reptation.f:2661-2727). */
int orc_eval(orc* o, const double* x, double* props, double* g, double* h,
             double* grad2_ph, int32_t* status);
int orc_eval_batch(orc* o, int nw, const double* x, double* props, double* g,
                   double* h, double* grad2_ph, int32_t* status);

/* The reference's 48-bit LCG rannyu (reptation.f:5247-5354). */
typedef struct orc_rng { int64_t l[4]; } orc_rng;
void orc_setrn(orc_rng* r, const int32_t seed[4]);
void orc_savern(const orc_rng* r, int32_t seed[4]);
double orc_rannyu(orc_rng* r);

/*
 * VMC replay: read_conf-style start from x0, then nsteps of move +
 * metropolis_test (reptation.f:4617-4680, 2622-2659) with the reference
 * stream seeded by seed[4] (state limbs as savern writes them).
 * bm_order selects which rannyu() of the Box-Muller line
 * (reptation.f:4638) is drawn first: 0 = the cos() argument, 1 = the log().
 * Per step writes p[s], age[s], ltf_new[s], e_new[s]; x_final gets x_old.
 */
int orc_vmc_replay(orc* o, const double* x0, int nsteps, double delta,
                   const int32_t seed[4], int bm_order, double* p, double* age,
                   double* ltf_new, double* e_new, double* x_final,
                   int32_t seed_final[4]);

#ifdef __cplusplus
}
#endif
#endif
