/*
 * peridem_gpu.h -- C-ABI of the MI355X explicit-step library (libperidem_gpu.so).
 *
 * The reference has no plugin API: its explicit step is the C++ template
 *   run_timeloop<dim>(PArr, TL, CN, Wall, CFGV)          (src/particle/timeloop.h:231-965)
 * called once from main() (run/simulate3d.cpp:142).  Its inputs are built by
 *   load_particles / load_contact / load_wall             (src/read/load_hdf5.h:41-181)
 * and its outputs are read back from ParticleN / RectWall every `modulo` steps
 *   (src/particle/timeloop.h:799-953).
 * This header replaces that template boundary with plain C entry points: the
 * host driver keeps the reference's config keys, time switches and output
 * layout, and hands the particle arrays across as plain pointers.
 *
 * Conventions
 *  - every call returns int: 0 on success, a negative PD_E* code on failure;
 *    pd_last_error(ctx) returns a message for the last failure.  Nothing
 *    exits the process and no exception crosses the boundary (the reference
 *    prints and exits: src/debug/err.h:21-24).
 *  - node arrays are row-major double[n*3] exactly as the HDF5 datasets
 *    (Pos, disp, vel, acc, extforce) are laid out; the library copies them.
 *  - a context owns one HIP stream on the current device; a context is used
 *    from one host thread at a time; distinct contexts are independent.
 *  - all arithmetic is IEEE FP64 without contraction; results are bitwise
 *    equal to a single-rank run of the reference algorithm.
 */
#ifndef PERIDEM_GPU_H
#define PERIDEM_GPU_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PD_OK 0
#define PD_E_ARG (-1)         /* bad argument / size mismatch */
#define PD_E_STATE (-2)       /* call not valid in this state (e.g. before pd_commit) */
#define PD_E_HIP (-3)         /* HIP runtime failure */
#define PD_E_UNSUPPORTED (-4) /* reference feature not built in this library */
#define PD_E_OVERFLOW (-5)    /* a fixed device capacity was exceeded (see message) */
#define PD_E_NOMEM (-6)       /* device memory budget exceeded */

/* Run-level parameters: Timeloop + Contact + RectWall<3> after apply_config
 * (src/particle/timeloop.h:71-229, src/particle/contact.h:104-183,185-376).
 * Wall planes are in the reference's 3-D lrtb order
 * [x_min, y_min, z_min, x_max, y_max, z_max] (contact.h:221-238). */
typedef struct pd_run_desc {
  double dt;
  uint32_t enable_fracture;       /* TL.enable_fracture */
  uint32_t self_contact;          /* CN.self_contact */
  uint32_t allow_damping;         /* CN.allow_damping */
  uint32_t allow_friction;        /* CN.allow_friction */
  uint32_t nl_bdry_only;          /* CN.nl_bdry_only */
  uint32_t compute_wall_reaction; /* CN.compute_wall_reaction */
  uint32_t gradient_extforce;     /* TL.gradient_extforce */
  uint32_t extforce_maxstep;      /* TL.extforce_maxstep */
  uint32_t enable_torque;         /* TL.enable_torque */
  uint32_t allow_wall;            /* Wall.allow_wall */
  double contact_rad;             /* CN.contact_rad */
  double normal_stiffness;        /* CN.normal_stiffness */
  double friction_coefficient;    /* CN.friction_coefficient */
  double damping_ratio;           /* CN.damping_ratio */
  double wall_lrtb[6];            /* Wall.lrtb() */
  double wall_speed[6];           /* Wall.speed_{x,y,z}_min, speed_{x,y,z}_max */
  uint32_t first_step;            /* value of t for the first step (1 for a fresh run) */
  uint32_t reserved;
} pd_run_desc;

/* One particle as load_particles reads it from P_%05u/ (load_hdf5.h:53-106). */
typedef struct pd_particle_desc {
  uint32_t nnodes;
  uint32_t instance;          /* independent system id (0 for a plain run) */
  const double* pos;          /* [nnodes*3] Pos */
  const double* vol;          /* [nnodes]   Vol */
  const double* disp;         /* [nnodes*3] disp (NULL = zeros) */
  const double* vel;          /* [nnodes*3] vel  (NULL = zeros) */
  const double* acc;          /* [nnodes*3] acc  (NULL = zeros) */
  const double* extforce;     /* [nnodes*3] extforce (NULL = zeros) */
  uint64_t nconn;             /* rows of Connectivity */
  const uint32_t* conn;       /* [nconn*2] Connectivity, file order */
  uint32_t nclamped;
  const uint32_t* clamped;    /* clamped_nodes */
  uint32_t nbdry;
  const uint32_t* bdry;       /* bdry_nodes (used only when nl_bdry_only) */
  double delta, rho, cnot, snot;
  int32_t movable, breakable, stoppable;
  uint32_t torque_axis;
  double torque_val;
} pd_particle_desc;

typedef struct pd_counters {
  uint64_t steps;              /* steps advanced since pd_commit */
  uint64_t bond_updates;       /* directed CSR entries of movable particles evaluated */
  uint64_t pair_computations;  /* particle pairs that passed the box test (timeloop.h:636-653) */
  uint64_t alive_bonds;        /* directed bonds not broken (computed on request) */
  uint64_t total_bonds;        /* directed bonds at commit */
  uint64_t nodes;
  uint64_t particles;
  uint64_t cell_rebuilds;      /* contact candidate rebuilds */
  uint64_t list_capacity;      /* contact candidates a node's list holds (grows on demand) */
  uint64_t list_regrows;       /* times the lists grew inside a pd_step */
} pd_counters;

typedef struct pd_ctx pd_ctx;

/* Create a context: replaces the value copies of Timeloop/Contact/RectWall
 * passed into run_timeloop (timeloop.h:231-233). */
int pd_create(pd_ctx** out, const pd_run_desc* run);

/* Add one particle: replaces load_particles' per-particle body
 * (load_hdf5.h:61-106) including conn2NArr (load_hdf5.h:27-37) and gen_xi
 * (particle2.h:241-257), which run at pd_commit.  Returns the particle index
 * through *index when index != NULL. */
int pd_add_particle(pd_ctx* ctx, const pd_particle_desc* p, uint32_t* index);

/* Build the neighbour rows in conn2NArr order, upload all state. */
int pd_commit(pd_ctx* ctx);

/* Advance nsteps explicit steps: the body of the loop at timeloop.h:395-797
 * (Verlet position half, PD force + bond breaking, self/wall/pair contact,
 * external force, Verlet velocity half, lagged boxes, wall motion).  The host
 * splits calls at time-switch and output boundaries. */
int pd_step(pd_ctx* ctx, uint32_t nsteps);

/* Host-side time switches (timeloop.h:405-535). */
#define PD_FLAG_ENABLE_FRACTURE 1
#define PD_FLAG_SELF_CONTACT 2
#define PD_FLAG_ALLOW_DAMPING 3
#define PD_FLAG_ALLOW_FRICTION 4
#define PD_FLAG_ENABLE_TORQUE 5
#define PD_FLAG_COMPUTE_WALL_REACTION 6
#define PD_FLAG_VELOCITY_CONSTRAINT 7   /* TL.enable_velocity_constraint (set_fix_velocity_timestep) */
int pd_set_flag(pd_ctx* ctx, int flag, int value);
int pd_set_wall_speed(pd_ctx* ctx, const double speed[6]);
int pd_set_wall(pd_ctx* ctx, const double lrtb[6]);
#define PD_PFLAG_MOVABLE 1
#define PD_PFLAG_STOPPABLE 2
#define PD_PFLAG_BREAKABLE 3
int pd_set_particle_flag(pd_ctx* ctx, uint32_t p, int which, int value);
/* TL.new_snot override (timeloop.h:314-318). */
int pd_set_snot(pd_ctx* ctx, uint32_t p, double snot);
/* Wheel reset (TL.reset_partzero_y at reset_partzero_y_timestep,
 * timeloop.h:496-535): call it before the pd_step whose first step is that
 * timestep.  Particle 0's reference positions move in y so that its mean
 * current position sits wheel_rad + contact_rad above the highest node of the
 * other particles (of its instance) lying within wheel_rad of it in x.  The
 * bond vectors keep the positions the particle was added with (gen_xi,
 * particle2.h:241-257); CurrPos, contact and walls see the moved ones.
 * wheel_rad is CFGV.wheel_rad (read_config.h:58,333-334).  Particle 1 must
 * have at least two nodes: the reference starts its running maximum from
 * particle 1's node 1. */
int pd_reset_partzero_y(pd_ctx* ctx, double wheel_rad);
/* Prescribed velocity of one particle (timeloop.h:685-764): the particle
 * CFGV.particle_w_fixed_velocity with fix_velocity_mean_x /
 * prescribed_velocity_mean_x and fix_velocity_angular /
 * prescribed_angular_vel.  Active while PD_FLAG_VELOCITY_CONSTRAINT is on;
 * pd_get_applied_force reads its applied_force_density [nnodes*3]
 * (zeros before the constraint has acted). */
int pd_set_velocity_constraint(pd_ctx* ctx, uint32_t p, int fix_mean_x, double mean_x,
                               int fix_angular, double angular_vel);
int pd_get_applied_force(pd_ctx* ctx, uint32_t p, double* afd);

/* State read-back for output (timeloop.h:826-851).  Any pointer may be NULL.
 * Arrays are [nnodes*3].  force is the force density, as in the reference. */
int pd_get_state(pd_ctx* ctx, uint32_t p, double* disp, double* currpos,
                 double* vel, double* acc, double* force);
/* Overwrite state (resume, timeloop.h:262-291).  Any pointer may be NULL.
 * A new disp also resets the particle's lagged box to the min/max of
 * pos + disp, as a resumed run starts (timeloop.h:389-393). */
int pd_set_state(pd_ctx* ctx, uint32_t p, const double* disp,
                 const double* vel, const double* acc);
/* Intact bonds as NbdArr2conn emits them (timeloop.h:20-44): pairs (i,q)
 * with i<q in row order.  *m receives the number of pairs; when cap is too
 * small nothing is written and PD_E_ARG is returned with *m set. */
int pd_get_connectivity(pd_ctx* ctx, uint32_t p, uint32_t* pairs,
                        uint64_t cap, uint64_t* m);
/* Wall planes (lrtb order) and, when compute_wall_reaction is on, the
 * reaction of the last step as get_reaction() orders it (contact.h:255-272):
 * [x_min, y_min, z_min, x_max, y_max, z_max] x 3 components. */
int pd_get_wall(pd_ctx* ctx, uint32_t job, double lrtb[6], double reaction[18]);
int pd_get_counters(pd_ctx* ctx, pd_counters* out);
/* Time of the last pd_step call on the device (HIP events), milliseconds. */
int pd_last_step_ms(pd_ctx* ctx, double* ms);

/* ---------------------------------------------------------------------------
 * Job pool: many independent systems advanced by the same launches.
 *
 * No counterpart in the reference, which runs one system per process
 * (run/simulate3d.cpp:142).  A job is what one reference run holds -- its
 * particles (pd_particle_desc, the `instance` field ignored), its
 * pd_run_desc -- plus a step budget.  Each pd_step advances every job that
 * has budget left by min(nsteps, remaining) steps, each job with its own dt,
 * clock t, walls and switches; a job whose budget is used up stays frozen
 * (state unchanged) until it is read back and removed.  A job's results are
 * bitwise those of the same system run alone.  pd_commit is the plain case:
 * one job per instance value, unlimited budget, plain particle indices
 * 0..n-1 in the order of pd_add_particle.
 *
 * Capacities are fixed by pd_reserve (one allocation of device memory);
 * pd_add_job fails with PD_E_NOMEM when a job does not fit in the free
 * space.  pd_job_footprint gives what a job needs, for admission control. */
typedef struct pd_pool_desc {
  uint64_t max_nodes;         /* node slots (rounded up to 64) */
  uint64_t max_bond_entries;  /* sliced bond entries (see pd_job_footprint) */
  uint64_t max_particles;
  uint64_t max_jobs;
} pd_pool_desc;
int pd_reserve(pd_ctx* ctx, const pd_pool_desc* pool);
/* node_slots, bond_entries: pool units the job occupies; bytes: device
 * memory they stand for.  Any output pointer may be NULL. */
int pd_job_footprint(const pd_particle_desc* parts, uint32_t nparts, uint64_t* node_slots,
                     uint64_t* bond_entries, uint64_t* bytes);
/* nsteps = UINT64_MAX: no budget limit.  *job receives the job slot. */
int pd_add_job(pd_ctx* ctx, const pd_run_desc* run, const pd_particle_desc* parts,
               uint32_t nparts, uint64_t nsteps, uint32_t* job);
int pd_remove_job(pd_ctx* ctx, uint32_t job);
/* pd_add_job in two halves, so that a server can lay out the next job on
 * the host while the pool steps: pd_job_prepare builds the rows, node arrays
 * and bond slices of a job without any context (thread-safe; errors are
 * read with pd_last_error(NULL) on the same thread); pd_add_prepared places
 * a prepared job into a pool (PD_E_NOMEM when it does not fit: the blob
 * stays valid and can be added later).  The caller frees blobs. */
typedef struct pd_job_blob pd_job_blob;
int pd_job_prepare(const pd_run_desc* run, const pd_particle_desc* parts, uint32_t nparts,
                   pd_job_blob** out);
int pd_job_blob_footprint(const pd_job_blob* blob, uint64_t* node_slots, uint64_t* bond_entries,
                          uint64_t* bytes);
int pd_add_prepared(pd_ctx* ctx, const pd_job_blob* blob, uint64_t nsteps, uint32_t* job);
void pd_job_blob_free(pd_job_blob* blob);
int pd_job_remaining(pd_ctx* ctx, uint32_t job, uint64_t* remaining);
/* pd_get_state / pd_get_connectivity / pd_set_flag for particle p of one
 * job; pd_get_wall takes the job slot as its instance argument. */
int pd_job_get_state(pd_ctx* ctx, uint32_t job, uint32_t p, double* disp, double* currpos,
                     double* vel, double* acc, double* force);
int pd_job_get_connectivity(pd_ctx* ctx, uint32_t job, uint32_t p, uint32_t* pairs,
                            uint64_t cap, uint64_t* m);
int pd_job_set_flag(pd_ctx* ctx, uint32_t job, int flag, int value);
/* Served mode: a request loop over a reserved pool, for C and C++ hosts.
 * No counterpart in the reference (one system per process,
 * run/simulate3d.cpp:142).  pd_serve_submit copies a request (run
 * parameters, particles, step budget) and lays it out on a worker thread;
 * each pd_serve_poll call admits waiting requests in arrival order while
 * their footprint fits the pool (pd_add_prepared), advances every resident
 * job by one chunk (at most `chunk` steps, never past the earliest budget end,
 * so finished jobs leave and waiting ones enter at call boundaries) and
 * returns the requests that finished.  A finished request's results stay in
 * its pool slot (pd_job_get_state, pd_job_get_connectivity, pd_get_wall with
 * that job) until pd_serve_release.  status is PD_OK, or why the request
 * produced no results: PD_E_NOMEM (larger than the empty pool), PD_E_OVERFLOW
 * (stopped by the library, see pd_step), a layout error.  Times are seconds
 * since pd_serve_create.  A server is used from one host thread; the pool
 * must not be stepped by anything else while it serves. */
typedef struct pd_server pd_server;
typedef struct pd_serve_event {
  uint64_t request;   /* id from pd_serve_submit */
  uint32_t job;       /* pool job slot holding the results (UINT32_MAX: none) */
  int32_t status;
  double t_submit, t_admit, t_done;
} pd_serve_event;
int pd_serve_create(pd_ctx* pool, uint32_t chunk, uint32_t prepare_threads, pd_server** out);
int pd_serve_submit(pd_server* server, const pd_run_desc* run, const pd_particle_desc* parts,
                    uint32_t nparts, uint64_t nsteps, uint64_t* request);
int pd_serve_poll(pd_server* server, pd_serve_event* events, uint32_t cap, uint32_t* n);
int pd_serve_pending(pd_server* server, uint64_t* queued, uint64_t* resident, uint64_t* unreported);
int pd_serve_release(pd_server* server, uint32_t job);
void pd_serve_destroy(pd_server* server);

/* Runs of 8 steps replay one captured HIP graph (default on); off: every
 * step is launched kernel by kernel.  Results are the same either way. */
int pd_set_graphs(pd_ctx* ctx, int on);

/* Bond-kernel shapes and their autotuner.  No counterpart in the reference.
 * The library carries several instantiations of the bond kernel (entries
 * per batch, software pipelining of the index and neighbour loads, and
 * "staged" shapes that first copy each 256-node block's table of referenced
 * nodes and their positions and volumes into LDS, and a shape with four
 * lanes per node, pipelined = 3, for pools too small to fill the GPU); each
 * sums every row in row order, so all give the same bits.
 * Staged shapes are usable while every resident job's block tables hold at
 * most 2240 nodes (pools of 2^27 node slots or more use one unstaged shape);
 * a selected shape that is not usable falls back to shape 4.  By default the
 * first pd_step of a pool with at least 2^18 resident nodes times every
 * usable shape on the resident state (scratch copies of the force and bond
 * arrays, so nothing changes) and keeps the fastest: every usable shape is
 * timed once (while staged shapes are usable, of the unstaged ones only shape
 * 4), and the shapes within 15% of the fastest three more times in
 * interleaved rounds, each keeping its smallest time (the four-lane shape
 * only on pools below 2^18 nodes); smaller pools use shape 14 (four lanes per
 * node) while fewer than 2^15 nodes are resident, else shape 10 (staged), or
 * 4 while that is not usable.  pd_autotune re-runs that with
 * reps timed launches per timing, returning the choice and (when ms != NULL)
 * the time of each shape [count], -1 for a shape not usable on this pool; pd_set_bond_shape forces one (-1: tune
 * at the next step); pd_get_bond_shape reports the shape launches use.
 * Shape index pd_bond_shape_count() (never tuned) is the unstaged kernel with
 * 64-bit record offsets that large pools use; it can be forced on any pool. */
/* Bond-term arithmetic.  No counterpart in the reference (a CPU code using
 * IEEE sqrt and division).  PD_ARITH_BITWISE (default): correctly rounded
 * square roots and divisions in the reference's operation order, results
 * bitwise the reference's.  PD_ARITH_TOLERANCE: |e| and |xi| come from the
 * hardware reciprocal square root refined by one Newton-Raphson step, the
 * divisions become multiplications by those reciprocals and the squared norms
 * and the row sums are contracted into fused multiply-adds, so every stretch is
 * within ~1e-14 (absolute) of the bitwise one: per step, each node's force
 * density agrees with the reference's to 1e-12 x c0 sum_j V_j over its row
 * (a stretch error of 1e-12), and the broken-bond set is exact for every
 * bond whose stretch lies farther than 1e-12 from s0.  Contact, wall,
 * external-force and Verlet arithmetic stay bitwise.  Changing the mode
 * re-tunes the bond shape at the next step (a forced shape is dropped);
 * pools of 2^27 node slots or more keep the bitwise kernel. */
#define PD_ARITH_BITWISE 0
#define PD_ARITH_TOLERANCE 1
int pd_set_arithmetic(pd_ctx* ctx, int mode);
int pd_get_arithmetic(pd_ctx* ctx, int* mode);

int pd_bond_shape_count(void);
int pd_bond_shape_info(int shape, int* unroll, int* pipelined, int* staged);
int pd_set_bond_shape(pd_ctx* ctx, int shape);
int pd_get_bond_shape(pd_ctx* ctx, int* shape);
int pd_autotune(pd_ctx* ctx, uint32_t reps, int* best, double* ms);

/* Per-phase device time, measured with HIP events on the context's stream
 * while profiling is on: profiled steps are launched one by one (not replayed
 * from a graph) with a mark before each phase (PD_PROFILE_PHASES) or only
 * around the bond kernel (PD_PROFILE_BOND: fewer marks in the stream, only
 * the bond phase is reported), and each step's marks are read eight steps
 * later, so the stream stays busy.  pd_set_profiling(ctx, 0) switches it off.
 * No counterpart in the reference, which only times whole modulo blocks
 * (timeloop.h:909-914). */
#define PD_PROFILE_PHASES 1
#define PD_PROFILE_BOND 2
#define PD_PHASE_POSITION 0 /* Verlet position half + boxes */
#define PD_PHASE_CONTACT_LIST 1 /* pair box test + cell-list rebuild */
#define PD_PHASE_BOND 2     /* bond force + fused break + the node's contact,
                               wall, external and torque terms + Verlet
                               velocity half (one kernel) */
#define PD_PHASE_CONTACT 3  /* wall reaction sum (steps that report it) */
#define PD_PHASE_UPDATE 4   /* clamped displacements, prescribed velocity,
                               lagged boxes, walls */
#define PD_NPHASE 5
typedef struct pd_profile {
  double ms[PD_NPHASE];
  uint64_t launches[PD_NPHASE];
} pd_profile;
int pd_set_profiling(pd_ctx* ctx, int on);
int pd_get_profile(pd_ctx* ctx, pd_profile* out); /* resets the accumulators */
const char* pd_last_error(pd_ctx* ctx);
void pd_destroy(pd_ctx* ctx);

/* The bond kernel evaluates sqrt and division with the backend's own
 * instruction sequences minus their operand-rescaling wrappers, valid while
 * coordinates, displacements and bond lengths stay in [2^-600, 2^600] (or
 * are 0); outside that domain it switches, for good, to the IEEE operators.
 * Results are bitwise the same either way.  pd_math_mode reports whether the
 * context has switched (*safe = 1).  pd_selftest_math checks the fast
 * sequences against IEEE `/` and `sqrt` on n random operands of the domain on
 * the current device and returns the mismatch counts (expected 0).  No
 * counterpart in the reference (a CPU code using the C library's sqrt). */
int pd_math_mode(pd_ctx* ctx, int* safe);
int pd_selftest_math(uint64_t n, uint64_t seed, uint64_t* mismatch_div, uint64_t* mismatch_sqrt);

/* pd_selftest_math without any device allocation, for callers that cannot
 * allocate: harnesses that forbid hipMalloc, serving processes that check the
 * arithmetic before admitting jobs.  No counterpart in the reference.  Same
 * kernel and counts as pd_selftest_math; *evaluated (may be NULL) receives
 * the number of samples checked (n).  The counters live in a 256-byte
 * file-scope device array of the library, one copy per GPU, held until the
 * process ends; every call on a GPU clears and reuses that GPU's copy, so
 * calls that overlap on one GPU race.  n above 2^40 is rejected (PD_E_ARG);
 * errors are read with pd_last_error(NULL) on the calling thread. */
int pd_selftest_math_noalloc(uint64_t n, uint64_t seed, uint64_t* mismatch_div,
                             uint64_t* mismatch_sqrt, uint64_t* evaluated);

/* Library version and build target, for logs. */
const char* pd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PERIDEM_GPU_H */
