/*
 * gsph.h - C ABI of the MI355X-native hybrid SPH step.
 *
 * The reference has no plugin or FFI layer: its hot path sits behind the
 * in-process C++ classes sph::HybridSystem (src/simulation/sph_hybrid_system.h:46-137)
 * and sph::Arrangement (src/grid/sph_arrangement.h:28-166) plus the solver
 * wrappers in src/solver/kernel_dispatch.cuh:22-28,67-83,106.  Each entry point
 * below replaces one of those seams; the comment on each names the reference
 * interface it stands in for.
 *
 * Conventions
 *   - every function returns an int status: GSPH_OK (0) or a negative GSPH_ERR_*;
 *     nothing in the library calls exit() (the reference's CUDA_SAFE_CALL does,
 *     src/core/cuda_call_check.h:15-39).
 *   - gsph_last_error(ctx) returns the message of the last failure on that ctx
 *     (gsph_last_error(NULL) the last failure that had no ctx).
 *   - host pointers are plain arrays in host memory; particle arrays are
 *     interleaved: pos4 = {x,y,z,w}*n with w = 1/density, vel3 = {x,y,z}*n,
 *     ev4 = {x,y,z,p}*n with p = pressure, acc3 = {x,y,z}*n.  This is the
 *     reference's ParticleBufferList SoA (src/particle/particle_buffer.h:116-172).
 *   - a ctx owns all of its device memory and one non-blocking HIP stream of
 *     default priority; calls on one ctx must be externally serialised, distinct
 *     ctx may be used from distinct threads.
 */
#ifndef GSPH_H_
#define GSPH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSPH_OK 0
#define GSPH_ERR_INVALID_ARG (-1)
#define GSPH_ERR_HIP (-2)
#define GSPH_ERR_OUT_OF_DOMAIN (-3)
#define GSPH_ERR_CAPACITY (-4)
#define GSPH_ERR_NO_DEVICE (-5)
#define GSPH_ERR_SPIN_TIMEOUT (-6)
#define GSPH_ERR_IO (-7)
#define GSPH_ERR_BUDGET (-8)
#define GSPH_ERR_INJECTED (-9)

/* Physics inputs of a scene.  The reference hard-codes these in
 * defaultInitializeSPHSysPara (src/simulation/sph_hybrid_system.cpp:127-202);
 * here they are inputs, with gsph_params_default() returning the reference's
 * values (h = 0.03, g = -9.8, dt = 0.003, ...). */
typedef struct gsph_params {
    float kernel_h;       /* smoothing length h; also the grid cell size */
    float mass;           /* particle mass (one global mass, as the reference) */
    float rest_density;
    float gas_constant;
    float viscosity;
    float time_step;
    float gravity[3];
    float wall_damping;
    float surface_normal; /* |grad colour| threshold for surface tension */
    float surface_coe;
    float world_size[3];  /* box extent; the grid is ceil(world/h) cells per axis */
} gsph_params;

/* Derived constants, the FP32 values the kernels use (the reference's
 * SystemParameter, src/core/sph_parameter.h:18-90, derived as in
 * sph_hybrid_system.cpp:143-193). */
typedef struct gsph_constants {
    float kernel_h, kernel_2, cell_size, inv_cell_size;
    int grid_size[3];
    float mass, rest_density, gas_constant, viscosity, time_step;
    float gravity[3];
    float wall_damping, surface_normal, surface_coe;
    float poly6_value, spiky_value, visco_value, grad_poly6, lplc_poly6;
    float self_density;
    float inv_rest_density, density_scale, half_spiky_value;
    float viscosity_visco_value, grad_color_scale, lplc_color_scale;
    float bound_min[3], bound_max[3];
    float world_size[3];
} gsph_constants;

/* 32-byte SMS task descriptor, the layout of the reference's BlockTask
 * (src/particle/particle_buffer.h:60-95). */
typedef struct gsph_task {
    int8_t is_same;
    uint8_t pad0;
    uint16_t paired_bounds; /* 2 bits each: minx,maxx,miny,maxy,minz,maxz */
    int32_t cellid;
    uint16_t cell_pos[3];
    uint16_t p_offset;
    int8_t xxi, xxx, yyi, yyy, zzi, zzz;
    int8_t pad1[2];
    int32_t cell_begin;
    int32_t cell_particle_count;
} gsph_task;

/* The reference's SmsTaskPairStats (src/grid/sph_arrangement.h:17-26). */
typedef struct gsph_pair_stats {
    int total_pairs;
    int same_cell_pairs;
    int is_same_pairs;
    int full_is_same_pairs;
    int partial_is_same_pairs;
    int compact_is_same_pairs;
    int widened_is_same_pairs;
} gsph_pair_stats;

/* Per-stage device times of one step in ms, from HIP events on the ctx stream
 * (the reference's tick_events_, sph_hybrid_system.cpp:266-333). */
typedef struct gsph_stage_ms {
    float arrange;
    float density;
    float force;
    float integrate;
    float total;
} gsph_stage_ms;

/* Scheduling options.  split_mode 0 = the reference split heuristic
 * (src/io/gpu_model.cu:172-180, bit-exact task assignment); 1 = tuned
 * thresholds below (same physics, different TRA/SMS assignment).  The
 * remaining fields are launch-shape knobs; gsph_autotune picks them by
 * timing the context's own scene.  Only split_mode/thresholds change which
 * path a particle takes; every option but same_cell_pair_force gives
 * bit-identical results (the TRA and SMS paths sum each particle's
 * neighbours in the same order).  same_cell_pair_force sums the pairs inside
 * an SMS cell in another order: its results agree with the others within the
 * FP32 rounding of the sums, not bit for bit, and gsph_autotune never turns
 * it on. */
typedef struct gsph_options {
    int split_mode;
    int sms_min_self;     /* tuned: cell is TRA when n <  sms_min_self ... */
    int sms_min_total;    /* ... and the 7-cell total < sms_min_total */
    int blocks_per_cu;    /* persistent grid size of density/force, per CU */
    int pair_masks;       /* 1: density records accepted pairs for force */
    int row_prune;        /* 1: skip rows/layers farther than h (exact superset) */
    int stage_depth;      /* LDS staging depth of TRA neighbour rows: 64, 128, 256 */
    int sms_mask_bytes;   /* SMS accepted-pair mask pool, bytes per particle of
                             capacity (set at creation; 0 = none) */
    int sms_local_merge;  /* 1: the two tasks of a judgeTask pair in one cell
                             stage their shared rows once, in density and force
                             (GSPH_ENABLE_SMS_LOCAL_MERGE and ..._FORCE,
                             CMakeLists.txt:76-91); 0: each task stages its own */
    int sms_legacy_staging; /* 1: the reference's legacy shared-memory SMS
                             staging (GSPH_USE_REGISTER_SMS=0,
                             kernel_common.cuh:282-519, 886-1151): isSame pairs
                             always share their rows, and dense rows below or
                             above the cell are read per micro-x layer, cut to
                             the task's micro-z bounds; same results */
    int same_cell_pair_force; /* 1: pairs inside an SMS cell of at most 512
                             particles are computed once by a pass of their
                             own and both sides accumulated
                             (GSPH_ENABLE_SAME_CELL_PAIR_FORCE,
                             force_kernels.cu:9-131, CMakeLists.txt:92-100);
                             needs 40 B per particle of capacity */
    int lean_passes;      /* 1: density and force first run lean passes that
                             take the TRA tiles needing no four-chain sums
                             and no pool rounds (force: with density's pair
                             masks), at higher occupancy; the general passes
                             take the rest.  0: the general passes take
                             every tile.  Same results */
} gsph_options;

typedef struct gsph_ctx gsph_ctx;

/* Library ------------------------------------------------------------------ */
const char *gsph_version(void);
int gsph_device_count(int *count);
const char *gsph_last_error(const gsph_ctx *ctx);
void gsph_params_default(gsph_params *out);
void gsph_options_default(gsph_options *out);
/* Replaces defaultInitializeSPHSysPara (sph_hybrid_system.cpp:127-202). */
int gsph_derive_constants(const gsph_params *p, gsph_constants *out);

/* Scenes (host only) ------------------------------------------------------- */
/* Reads the reference scene JSON (assets/scene_default.json format; reader at
 * sph_hybrid_system.cpp:69-126) and generates its lattice exactly as
 * initializeScene does (sph_hybrid_system.cpp:432-473).  Optional extension
 * keys: kernel_h, gravity_y, time_step, rest_density.  Call with pos4 = NULL
 * to get the count, then again with a buffer of 4*count floats. */
int gsph_scene_json(const char *path, gsph_params *params, int *count,
                    float *pos4, int capacity);
/* Synthetic workloads: 0 = 32^3 dam-break block, 1 = 158^3 column in a
 * 6.32 m cube, 2 = pool plus fine sphere.  Same two-call protocol. */
int gsph_scene_config(int config, gsph_params *params, int *count,
                      float *pos4, int capacity);
/* One jittered lattice block (the building block of the synthetic scenes):
 * nx*ny*nz particles at corner + (i+1/2)*spacing, jitter +-jitter_frac*spacing
 * from SplitMix64(seed). */
int gsph_scene_block(const float corner[3], int nx, int ny, int nz,
                     float spacing, float jitter_frac, uint64_t seed,
                     float *pos4);

/* Batched form --------------------------------------------------------------- */
/* No counterpart in the reference (it runs one scene per process).  nscenes
 * independent scenes share one context: scene s has box world3[3s..3s+2] and
 * counts[s] particles, stored back to back in pos4 / vel3 (vel3 may be NULL).
 * The physics constants (h, mass, rho0, dt, ...) are shared; p->world_size is
 * ignored.  Every gsph_* step call then advances all scenes with one launch
 * per kernel.  Scenes never interact, and each scene's particles keep the
 * slot range [first, first + count) given by gsph_scene_range. */
int gsph_batch_create(const gsph_params *p, const gsph_options *opt, int nscenes,
                      const float *world3, const int *counts, const float *pos4,
                      const float *vel3, gsph_ctx **out);
int gsph_scene_count(const gsph_ctx *ctx, int *nscenes);
int gsph_scene_range(const gsph_ctx *ctx, int scene, int *first, int *count);
/* Box of a scene as given at creation (plain and batch contexts). */
int gsph_scene_world(const gsph_ctx *ctx, int scene, float world[3]);
/* Parameters a context was created with; whether it is a batch context. */
int gsph_get_params(const gsph_ctx *ctx, gsph_params *out);
int gsph_is_batched(const gsph_ctx *ctx, int *batched);
/* Completed steps (integrations) of a context; settable for restarts. */
int gsph_step_count(const gsph_ctx *ctx, long long *steps);
int gsph_set_step_count(gsph_ctx *ctx, long long steps);

/* Checkpoint / resume.  No counterpart in the reference, whose state only
 * moves between device and host buffers (ParticleBufferObject::transfer,
 * src/particle/particle_buffer.cpp:369-381).  A snapshot file holds the
 * parameters, options, scene boxes and counts, the step count and the state
 * (pos4, vel3, ev4) in the context's current slot order, with an FNV-1a 64
 * checksum; a context loaded from it continues bit-exactly.  The file is
 * written under path + ".tmp" and renamed.  gsph_snapshot_check validates a
 * file without touching the GPU.  opt (nullable) overrides the stored
 * options on load. */
int gsph_snapshot_save(gsph_ctx *ctx, const char *path);
int gsph_snapshot_check(const char *path, long long *n, int *nscenes, long long *steps);
int gsph_snapshot_load(const char *path, const gsph_options *opt, gsph_ctx **out);

/* Serving form --------------------------------------------------------------- */
/* No counterpart in the reference.  A continuous-batching request loop:
 * requests (one scene + a step count) are admitted in arrival order against
 * an HBM budget (exact engine bytes, gsph_engine_bytes), a scene-slot limit
 * and the 2^26-cell grid limit (a request that does not fit yet waits); all
 * resident scenes advance together one step per engine iteration with no host
 * synchronisation per step (at most 3 steps in flight, device error flags
 * read one step behind); the layout changes only when scenes join or leave,
 * and then without a host synchronisation (continuing scenes keep their
 * device state and sorted order); finished scenes are copied out
 * asynchronously into pinned host memory.  A request larger than the budget,
 * or with particles outside its box, is rejected at submission; a scene whose
 * state leaves its box or stops being finite fails alone.  run(realtime = 1)
 * takes arrival times in seconds against the wall clock; realtime = 0 counts
 * them in engine iterations. */
typedef struct gsph_server gsph_server;
typedef struct gsph_server_stats {
    int completed, failed, rejected, max_active;
    double p50, p99;          /* request latency (arrival -> result), s or iterations */
    double wall;              /* s */
    double particle_steps;
    long long engine_steps, recompositions;
    double peak_bytes;        /* admitted HBM estimate */
} gsph_server_stats;
#define GSPH_REQ_QUEUED 0
#define GSPH_REQ_ACTIVE 1
#define GSPH_REQ_DONE 2
#define GSPH_REQ_FAILED 3
#define GSPH_REQ_REJECTED 4
/* Device bytes of a context whose buffers hold `particles` particles, `cells`
 * grid cells and `scenes` scenes under options opt (NULL: defaults): exactly
 * what the library allocates for those capacities (every context checks its
 * allocation against it).  The serving loop charges its engine by this,
 * including both engines while a grown one replaces the old; a request is
 * rejected when an engine sized for it alone exceeds the budget.  Host only. */
int gsph_engine_bytes(const gsph_options *opt, long long particles, long long cells, int scenes,
                      double *bytes);
int gsph_server_create(const gsph_params *p, const gsph_options *opt,
                       double hbm_budget_bytes, int max_scenes, gsph_server **out);
int gsph_server_destroy(gsph_server *s);
int gsph_server_submit(gsph_server *s, const float world3[3], int n,
                       const float *pos4, const float *vel3, int steps,
                       double arrival, int *id);
int gsph_server_run(gsph_server *s, int realtime);
int gsph_server_result(gsph_server *s, int id, float *pos4, float *vel3,
                       float *ev4, int *state, double *latency, double *queued);
int gsph_server_get_stats(gsph_server *s, gsph_server_stats *out);
const char *gsph_server_last_error(const gsph_server *s);

/* Plain single scene -------------------------------------------------------- */
/* HybridSystem ctor + initializeScene + BuffInit (sph_hybrid_system.cpp:209-240,
 * 432-473; kernel_dispatch.cu:17-22): uploads n particles, ev = acc = 0. */
int gsph_create(const gsph_params *p, const gsph_options *opt, int n,
                const float *pos4, const float *vel3, gsph_ctx **out);
int gsph_destroy(gsph_ctx *ctx);
int gsph_num_particles(const gsph_ctx *ctx, int *n);
int gsph_get_constants(const gsph_ctx *ctx, gsph_constants *out);
/* Overwrites the device state (any of vel3/ev4 may be NULL = zero). */
int gsph_set_state(gsph_ctx *ctx, const float *pos4, const float *vel3,
                   const float *ev4);
/* HybridSystem::addParticle (src/simulation/sph_hybrid_system.cpp:602-627)
 * with Arrangement::resetNumParticle (src/grid/sph_arrangement.cu:1643-1670):
 * appends count particles to a plain context between steps.  pos4 gives x, y,
 * z (w is ignored and stored as 0, as addParticle does), vel3 may be NULL
 * (zero velocity), ev and acc start at 0.  Positions must be finite and
 * inside the box [0, world).  When the particle capacity is exceeded every
 * per-particle buffer grows to max(2 x capacity, n + count), as the
 * reference's buffers double, and the state is kept; unlike the reference,
 * the sort, split and task buffers grow with it.  reset_all != 0 also zeroes
 * ev and acc of the particles already there, as the reference's emitter
 * (action1, sph_hybrid_system.cpp:653-676) does with BuffInit over all
 * particles.  The next arrange sorts every particle; a captured step graph is
 * dropped. */
int gsph_add_particles(gsph_ctx *ctx, int count, const float *pos4, const float *vel3,
                       int reset_all);
/* Arrangement::arrangeHybridFrame (src/grid/sph_arrangement.cu:1493-1506):
 * hash, stable sort, permute, cell begin/count, split, compact index, tasks,
 * pairing.  No host synchronisation. */
int gsph_arrange(gsph_ctx *ctx);
/* computeDensityHybrid (src/solver/kernel_dispatch.cu:206-222). */
int gsph_density(gsph_ctx *ctx);
/* computeForceHybrid (src/solver/kernel_dispatch.cu:224-253). */
int gsph_force(gsph_ctx *ctx);
/* advance -> knIntegrateVelocityE (kernel_dispatch.cu:314-323). */
int gsph_integrate(gsph_ctx *ctx);
/* HybridSystem::tick x nsteps (sph_hybrid_system.cpp:253-339).  When
 * per_step_ms is non-NULL it receives nsteps stage records (events per step);
 * otherwise the steps run back to back with no host synchronisation. */
int gsph_step(gsph_ctx *ctx, int nsteps, gsph_stage_ms *per_step_ms);
/* Captures one full step as a hipGraph; later gsph_step calls replay it. */
int gsph_enable_graph(gsph_ctx *ctx, int enable);
/* Autotuner (replaces the reference's offline performance-model tables,
 * src/io/gpu_model.cu:137-181 and its launch constants).  Times trial_steps
 * steps of the context's own scene per candidate: persistent grid size
 * (blocks_per_cu 4/8/16/32) x LDS staging depth (64/128/256) x pair masks, then,
 * when tune_split is nonzero, the reference split rule against tuned
 * thresholds.  The state (pos, vel, ev) is restored after every trial (the
 * acceleration buffer keeps the last trial's values); the fastest options
 * stay set on ctx and are returned in *chosen (best_ms: its step time). */
int gsph_autotune(gsph_ctx *ctx, int trial_steps, int tune_split,
                  gsph_options *chosen, float *best_ms);
/* Options of a live context (validated; drops a captured graph). */
int gsph_set_options(gsph_ctx *ctx, const gsph_options *opt);
int gsph_get_options(gsph_ctx *ctx, gsph_options *out);
int gsph_sync(gsph_ctx *ctx);
/* Device -> host copies of the current state (sorted order).  NULL skips. */
int gsph_read_state(gsph_ctx *ctx, float *pos4, float *vel3, float *ev4,
                    float *acc3);
/* Sizes of the grid outputs: cells = C, tasks = T (SMS task count),
 * middle_raw = the reference's d_middle_value_ (-1 when there is no SMS
 * particle), tra = TRA particle count (middle clamped to [0, n]). */
int gsph_grid_info(gsph_ctx *ctx, int *cells, int *tasks, int *middle_raw,
                   int *tra);
/* Arrangement getters (sph_arrangement.h:41-57): sorted micro hash[n], the
 * stable permutation perm[n] (sorted slot -> previous slot), cell begin[C] and
 * count[C] (getDevCellOffset/getDevCellNumP), SMS tasks per cell req[C]
 * (d_block_reqs_), compact[n] (getDevCellIndex), tasks[T+1] (getSmsTasks,
 * including the odd-tail pad).  NULL skips an array. */
int gsph_read_grid(gsph_ctx *ctx, int *hash, int *perm, int *begin,
                   int *count, int *req, int *compact, gsph_task *tasks);
/* computeBenchmarkStateChecksum (sph_hybrid_system.cpp:43-62): sum x, y, z,
 * w and sum ((i&1023)+1)(x+3y+7z) in FP64, reduced on the device in a fixed
 * order. */
int gsph_checksum(gsph_ctx *ctx, double out[5]);
/* The same checksum over state slots [first, first + count), with i counted
 * from first (the checksum of one scene of a batch). */
int gsph_checksum_range(gsph_ctx *ctx, int first, int count, double out[5]);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation): the checksum of gsph_checksum_range for nranges ranges
 * [first[k], first[k] + count[k]) of the current state (say every scene of a
 * batch) in one call, bit-identical to gsph_checksum_range per range.  The
 * results go to out5_device, a device buffer of 5 * nranges doubles the caller
 * provides; first and count are host arrays.  Nothing is allocated and the
 * host does not wait: the kernels are enqueued on the context's stream (read
 * out5_device after gsph_sync).  The per-block partial sums live in a static
 * device array of the library (64 ranges x 256 blocks x 5 doubles, 640 KiB per
 * GPU, held until the process ends), so nranges <= 64 (GSPH_ERR_CAPACITY
 * beyond).  Every call on one GPU shares that array: calls that overlap on
 * one GPU (different contexts or streams) race; each call overwrites all the
 * partials it reads, so nothing carries over from an earlier call. */
int gsph_checksum_ranges_noalloc(gsph_ctx *ctx, int nranges, const int *first, const int *count,
                                 double *out5_device);
/* Arrangement::getSmsTaskPairStats (sph_arrangement.cu:1595-1636). */
int gsph_get_pair_stats(gsph_ctx *ctx, gsph_pair_stats *out);
/* Accepted-pair mask pool: rows (64 words each) taken by the last density
 * pass (SMS items and TRA tiles with rows longer than one mask word) and the
 * capacity (sms_mask_bytes per particle).  No counterpart in the reference,
 * which has no accepted-pair masks; a diagnostic for tuning sms_mask_bytes. */
int gsph_mask_pool_info(gsph_ctx *ctx, int *used_rows, int *capacity_rows);
/* Per-kernel device time (ms, summed) accumulated while kernel timing is on;
 * names are written as a ';'-separated list. */
int gsph_kernel_timing(gsph_ctx *ctx, int enable);
int gsph_kernel_times(gsph_ctx *ctx, char *names, int names_cap, float *ms,
                      int *launches, int cap, int *count);
/* Test hook: the next call that launches kernels fails with GSPH_ERR_INJECTED
 * after launching nothing (failure-path tests of the serving loop). */
int gsph_inject_fault(gsph_ctx *ctx, int after_calls);

#ifdef __cplusplus
}
#endif

#endif /* GSPH_H_ */
