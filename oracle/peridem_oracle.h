/*
 * peridem_oracle.h -- CPU oracle for the PeriDEM explicit step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (src/particle/particle2.h, src/particle/contact.h,
 * src/particle/timeloop.h, src/compat/overloads.h, src/read/load_hdf5.h),
 * used to check the HIP library's results and as the CPU baseline in the
 * benchmark.  Product code never links or calls it.
 *
 * The API mirrors include/peridem_gpu.h (or_* for pd_*) and takes the same
 * descriptor structs, so a test builds one system description and feeds it
 * to both.
 */
#ifndef PERIDEM_ORACLE_H
#define PERIDEM_ORACLE_H
#include "../include/peridem_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

int or_create(or_ctx** out, const pd_run_desc* run);
int or_add_particle(or_ctx* ctx, const pd_particle_desc* p, uint32_t* index);
int or_commit(or_ctx* ctx);
int or_step(or_ctx* ctx, uint32_t nsteps);
/* Emulate the reference's MPI decomposition with `nranks` OpenMP threads:
 * particle i and pair p go to rank i%nranks / p%nranks, each rank owns a
 * zeroed force buffer, buffers are summed in rank order (timeloop.h:544,
 * 637, 658-659).  nranks=1 (default) is the bitwise reference order. */
int or_set_ranks(or_ctx* ctx, int nranks);
/* Contact candidates from a cell grid over the current positions instead of
 * every node pair, visited in the full loop's order: bitwise the same sums,
 * far fewer distance tests on large dense systems.  Off by default (the
 * reference's all-pairs loop, which the CPU baseline times). */
int or_set_prune(or_ctx* ctx, int on);
int or_set_flag(or_ctx* ctx, int flag, int value);
int or_set_wall_speed(or_ctx* ctx, const double speed[6]);
int or_set_wall(or_ctx* ctx, const double lrtb[6]);
int or_set_particle_flag(or_ctx* ctx, uint32_t p, int which, int value);
int or_set_snot(or_ctx* ctx, uint32_t p, double snot);
/* Wheel reset of particle 0's reference positions (timeloop.h:496-535),
 * called before the step it precedes. */
int or_reset_partzero_y(or_ctx* ctx, double wheel_rad);
int or_get_state(or_ctx* ctx, uint32_t p, double* disp, double* currpos,
                 double* vel, double* acc, double* force);
int or_set_state(or_ctx* ctx, uint32_t p, const double* disp,
                 const double* vel, const double* acc);
int or_get_connectivity(or_ctx* ctx, uint32_t p, uint32_t* pairs,
                        uint64_t cap, uint64_t* m);
/* Last step's stretch of every bond row entry of particle p, in row order
 * (the reference's ParticleN::stretch, particle2.h:153); *m = entries. */
int or_get_stretch(or_ctx* ctx, uint32_t p, double* out, uint64_t cap, uint64_t* m);
int or_get_wall(or_ctx* ctx, uint32_t instance, double lrtb[6], double reaction[18]);
int or_set_velocity_constraint(or_ctx* ctx, uint32_t p, int fix_mean_x, double mean_x,
                               int fix_angular, double angular_vel);
int or_get_applied_force(or_ctx* ctx, uint32_t p, double* afd);
int or_get_counters(or_ctx* ctx, pd_counters* out);
const char* or_last_error(or_ctx* ctx);
void or_destroy(or_ctx* ctx);

/* Closed-form helper exported for tests: segment_mass<3> (contact.h:14-39)
 * returns centroid in out[0] and area in out[1]. */
void or_segment_mass3(double R, double d, double out[2]);

#ifdef __cplusplus
}
#endif
#endif
