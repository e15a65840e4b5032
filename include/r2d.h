/*
 * resolve2d on MI355X -- C ABI of the GPU step.
 *
 * This is the drop-in boundary for resolve2d's step call.  Each entry point
 * names the reference interface it replaces (paths relative to the resolve2d
 * source tree).  Conventions:
 *   - every call returns an r2d_status; 0 is success, so the reference's
 *     "bool false on error" of src/wasm_root.zig is `status != 0`;
 *   - worlds are opaque handles (the reference's wasm layer keeps one global
 *     `?*Solver`, src/wasm_root.zig:15); many worlds live in one context and
 *     can be stepped together;
 *   - bodies are addressed by stable ids exactly like RigidBody.Id
 *     (src/core/Bodies/RigidBody.zig:15), widened from u16 to u32 so worlds
 *     may exceed 65,535 bodies; ids are never reused;
 *   - iteration order (src/core/lib.zig:135, an AutoArrayHashMap in insertion
 *     order, with swapRemove on removal) is preserved and exposed through
 *     r2d_body_id_at, as src/wasm_root.zig:55-58 does;
 *   - one host thread per context; device work is ordered on the context's
 *     own HIP stream; body state stays resident in device memory and is
 *     copied only when the host reads or edits it.
 */
#ifndef R2D_H
#define R2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum r2d_status {
    R2D_OK = 0,
    R2D_E_INVALID_ARG = 1,   /* null handle, bad enum, zero-sized shape ...            */
    R2D_E_INVALID_ID = 2,    /* error.InvalidRigidBodyId (Constraints/DistanceJoint.zig:44-47 etc.),
                                error.NoSuchIdExists (src/core/lib.zig:309)               */
    R2D_E_OOM = 3,           /* error.OutOfMemory (host or device allocation)          */
    R2D_E_NONFINITE = 4,     /* a body's AABB is not finite: the reference's @intFromFloat
                                in SpatialHash.zig:93-96 is undefined there; the world is
                                left unstepped                                           */
    R2D_E_DEVICE = 5,        /* HIP runtime failure                                      */
    R2D_E_NO_DEVICE = 6,     /* no usable GPU                                            */
    R2D_E_RANGE = 7,         /* index out of range (r2d_body_id_at), too many cells     */
    R2D_E_CAPACITY = 8,      /* request larger than a fixed limit                        */
    R2D_E_BUSY = 9           /* the world is held by a server (r2d_serve_submit) until its
                                request completes; the context already has a server      */
} r2d_status;

typedef struct r2d_ctx r2d_ctx;
typedef struct r2d_world r2d_world;

/* EntityFactory.BodyOptions (src/core/lib.zig:35-46) */
typedef struct r2d_body_options {
    float pos_x, pos_y;
    float vel_x, vel_y;      /* default 0 */
    float angle;             /* default 0 */
    float omega;             /* default 0 */
    float mu;                /* default 0.5 */
    int32_t mass_kind;       /* 0: mass_value is a density (kg/m^2), 1: a mass (kg) */
    float mass_value;
} r2d_body_options;

/* Constraint.Parameters (src/core/Constraints/Constraint.zig:27-31):
 * defaults power_max = +inf, power_min = -inf, beta = 10 */
typedef struct r2d_constraint_params {
    float power_max, power_min, beta;
} r2d_constraint_params;

/* Everything src/wasm_root.zig's getters expose for one body (:88-251). */
typedef struct r2d_body_state {
    uint32_t id;
    uint32_t type;           /* 0 disc, 1 rectangle (RigidBody.Type) */
    uint32_t is_static;
    uint32_t num_normals;    /* 1 disc, 4 rectangle */
    float pos_x, pos_y;
    float mom_x, mom_y;
    float force_x, force_y;
    float mass;
    float angle;
    float ang_momentum;
    float torque;
    float inertia;
    float mu;
    float aabb_x, aabb_y, aabb_half_width, aabb_half_height;
    float radius;            /* discs */
    float width, height;     /* rectangles */
} r2d_body_state;

/* One contact manifold as src/core/collision.zig:54-71 stores it. */
typedef struct r2d_manifold_point {
    uint32_t valid;
    float acc_pn, acc_pt;
    float ref_r_x, ref_r_y, inc_r_x, inc_r_y;
    float pos_x, pos_y;
    float depth, original_depth;
    float mass_n, mass_t;
} r2d_manifold_point;

typedef struct r2d_manifold_state {
    uint32_t ref_id, inc_id;
    float normal_x, normal_y, tangent_x, tangent_y;
    float friction;
    r2d_manifold_point points[2];
} r2d_manifold_state;

/* --------------------------------------------------------------- context */
/* Creates a context on HIP device `device` with its own stream. */
int r2d_ctx_create(int device, r2d_ctx** out);
/* Destroy order: servers (r2d_serve_destroy), then worlds
 * (r2d_world_destroy), then the context.  R2D_E_BUSY while a server is
 * attached or a world created on the context is not yet destroyed. */
int r2d_ctx_destroy(r2d_ctx* ctx);
/* Waits for all device work of the context. */
int r2d_ctx_sync(r2d_ctx* ctx);

/* ----------------------------------------------------------------- world */
/* Solver.init (src/core/lib.zig:145-157).  Like the reference, the two hash
 * arguments are stored but the step uses cell width 4.0 and 2*N buckets
 * (lib.zig:254-255) unless r2d_set_hash_params selects another mode. */
int r2d_world_create(r2d_ctx* ctx, float spatialhash_cell_width, size_t spatialhash_table_size_mult,
                     r2d_world** out);
/* Solver.deinit (src/core/lib.zig:159-179). */
int r2d_world_destroy(r2d_world* w);
/* Solver.clear (src/core/lib.zig:181-187): drops bodies, constraints, force
 * generators, manifolds and excluded pairs; the id counter keeps counting. */
int r2d_world_clear(r2d_world* w);

/* EntityFactory.makeDiscBody (src/core/lib.zig:73-82). */
int r2d_make_disc(r2d_world* w, const r2d_body_options* bo, float radius, uint32_t* id_out);
/* EntityFactory.makeRectangleBody (src/core/lib.zig:84-93). */
int r2d_make_rect(r2d_world* w, const r2d_body_options* bo, float width, float height, uint32_t* id_out);
/* `body_unwrap().static = flag` (src/examples/0_1_car_platformer.zig:25, demos/native/src/main.zig). */
int r2d_set_static(r2d_world* w, uint32_t id, int is_static);
/* Solver.removeRigidBody (src/core/lib.zig:308-310): swapRemove, the last body
 * takes the removed body's iteration slot.  R2D_E_INVALID_ID if absent. */
int r2d_remove_body(r2d_world* w, uint32_t id);

/* EntityFactory.makeDownwardsGravity (src/core/lib.zig:95-98). */
int r2d_make_gravity(r2d_world* w, float g);
/* Anchored spring force generator on one body: F += -k (x - anchor) - c v.
 * Extension: the reference's ForceGen.Type has only downwards_gravity
 * (src/core/Forces/ForceGen.zig:12-14); this fits its per-body apply(). */
int r2d_make_spring(r2d_world* w, uint32_t id, float k, float c, float anchor_x, float anchor_y);

/* EntityFactory.makeDistanceJoint (src/core/lib.zig:106-110). */
int r2d_make_distance_joint(r2d_world* w, const r2d_constraint_params* p, uint32_t id1, uint32_t id2,
                            float target_distance, size_t* handle_out);
/* EntityFactory.makeOffsetDistanceJoint (src/core/lib.zig:100-104). */
int r2d_make_offset_distance_joint(r2d_world* w, const r2d_constraint_params* p, uint32_t id1, uint32_t id2,
                                   float r1_x, float r1_y, float r2_x, float r2_y, float target_distance,
                                   size_t* handle_out);
/* EntityFactory.makeFixedPositionJoint (src/core/lib.zig:112-116). */
int r2d_make_fixed_position_joint(r2d_world* w, const r2d_constraint_params* p, uint32_t id, float target_x,
                                  float target_y, size_t* handle_out);
/* EntityFactory.makeMotorJoint (src/core/lib.zig:118-122). */
int r2d_make_motor_joint(r2d_world* w, const r2d_constraint_params* p, uint32_t id, float target_omega,
                         size_t* handle_out);
/* EntityFactory.excludeCollisionPair (src/core/lib.zig:124-129). */
int r2d_exclude_collision_pair(r2d_world* w, uint32_t id1, uint32_t id2);

/* ------------------------------------------------------------------ step */
/* Solver.process (src/core/lib.zig:189-251), the hot path: broadphase and
 * narrowphase once, then sub_steps x (forces, integration, pre-step,
 * collision_iters x (constraints, manifolds), position update).
 * R2D_E_INVALID_ID if a constraint names a removed body (the reference
 * raises the same error from Constraint.solve; here the world is left
 * unstepped). */
int r2d_process(r2d_world* w, float dt, size_t sub_steps, size_t collision_iters);

/* Steps n independent worlds of one context together, `steps` times.
 * Equivalent to calling r2d_process on every world `steps` times.  A batch
 * whose worlds all have at most 1,024 bodies, the reference order and hash
 * tables that fit LDS is one device launch for all steps of all worlds (one
 * workgroup per world, r2d_ctx_set_fused); other batches take one launch
 * sequence per step (replayed as a hipGraph when it repeats).  The first
 * failing world's status is returned; other worlds are stepped regardless. */
int r2d_process_batch(r2d_ctx* ctx, r2d_world* const* worlds, size_t n, float dt, size_t sub_steps,
                      size_t collision_iters, size_t steps);

/* The step of r2d_process without any allocation, for callers that cannot
 * allocate device memory (code under hipGraph capture, serving loops and
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
 * step scratch is a static device array of 8 MiB that each GPU holds for
 * the life of the process, shared by every call on that GPU: calls that
 * overlap on one GPU race (calls from threads of this process are
 * serialised by a per-device lock).  Bit-identical to r2d_process.  The
 * world must be resident -- stepped once by r2d_process or
 * r2d_process_batch since bodies, joints or force generators were last
 * added or removed (r2d_set_body edits are fine) -- else R2D_E_INVALID_ARG;
 * at most 1,024 bodies, 4,096 hash buckets and 2,048 joints, the reference
 * order and no warm starting, else R2D_E_CAPACITY / R2D_E_INVALID_ARG.  A
 * step that would need more than 16,384 cell entries, 65,536 candidate
 * pairs, 4,096 manifolds or 6,144 solver items returns R2D_E_CAPACITY with
 * the world unchanged.  r2d_get_manifolds reads this step's manifolds until
 * the next r2d_process_static call on the same GPU (then R2D_E_RANGE). */
int r2d_process_static(r2d_world* w, float dt, size_t sub_steps, size_t collision_iters);

/* ----------------------------------------------------------- state I/O */
/* solverGetNumBodies (src/wasm_root.zig:51-53). */
int r2d_num_bodies(r2d_world* w, size_t* n_out);
/* solverGetBodyIdBasedOnIter (src/wasm_root.zig:55-58). */
int r2d_body_id_at(r2d_world* w, size_t iter_index, uint32_t* id_out);
/* Getters of src/wasm_root.zig:88-251 for one body. */
int r2d_get_body(r2d_world* w, uint32_t id, r2d_body_state* out);
/* Setters of src/wasm_root.zig:156-224 (momentum, force, angular momentum,
 * torque) plus position and angle, which callers write through props
 * (demos/native/src/main.zig:114-135).  Shape, mass and id are ignored. */
int r2d_set_body(r2d_world* w, uint32_t id, const r2d_body_state* in);
/* All bodies in iteration order; *n_out = count (written even if cap is short). */
int r2d_get_bodies(r2d_world* w, r2d_body_state* out, size_t cap, size_t* n_out);
/* The manifolds of the last step in the reference's order (Solver.manifolds,
 * read by demos/native/src/Renderer.zig:84-167). */
int r2d_get_manifolds(r2d_world* w, r2d_manifold_state* out, size_t cap, size_t* n_out);

/* ------------------------------------------------------------ options */
/* Spatial-hash parameters.  mode 0 (default): the reference's hard-coded
 * cell 4.0 and 2*N buckets (src/core/lib.zig:254-255), bit-exact pair set.
 * mode 1: the given cell width and table_mult*N buckets (the roadmap's
 * "wire solver spatial-hash configuration", README.md:62-63).  Discovery
 * order, and so results, can differ from mode 0. */
int r2d_set_hash_params(r2d_world* w, int mode, float cell_width, size_t table_mult);
/* Solver schedule.  0 (default): the reference's sequential Gauss-Seidel
 * order, reproduced exactly.  1: graph-coloured batches (same equations,
 * another order; results agree within tolerance, not bit for bit).
 * Schedule 1 is refused (R2D_E_INVALID_ARG) while warm starting is on. */
int r2d_set_schedule(r2d_world* w, int mode);
/* Warm starting (opt-in; the roadmap's "persist manifolds across frames by
 * keying with stable body IDs", README.md:60-61 -- the reference rebuilds
 * every manifold with zero impulses, lib.zig:261 and collision.zig:38-51).
 * factor 0 (default): off, the reference's results.  factor in (0, 1]:
 * after each step the world keeps every manifold (body ids, points, normal,
 * accumulated impulses) under the lower slot of its two bodies, at most 8
 * per body in discovery order.  A manifold of the next step with the same
 * (reference id, incident id), the same point slots, a normal within
 * cos = 0.99 of the kept one and no point that updateTGSDepth moves starts
 * every sub-step with factor * impulse / sub_steps per point: added to its
 * accumulators before the iterations and applied to its bodies with the
 * forces (each body's share summed in item order).  Results then differ
 * from the reference's by design.  Removing a body, clearing the world or
 * calling this function empties the kept manifolds.  Needs schedule 0. */
int r2d_set_warm_start(r2d_world* w, float factor);

/* --------------------------------------------------------- profiling */
/* When on, every step records HIP events around its phases on the
 * context's stream (broadphase, narrowphase, scheduling, solve). */
int r2d_ctx_set_profiling(r2d_ctx* ctx, int on);
/* Accumulated milliseconds and step count per phase: out_ms[4], *steps. */
int r2d_ctx_phase_times(r2d_ctx* ctx, double* out_ms, uint64_t* steps);
/* Device time (ms) of each timed step since the last reset, in order: up to
 * cap values into out (may be NULL), the count into *n.  Steps taken by the
 * fused kernel report each world's step (wall clock of its workgroup). */
int r2d_ctx_step_times(r2d_ctx* ctx, float* out, size_t cap, size_t* n);
int r2d_ctx_reset_phase_times(r2d_ctx* ctx);
/* Profiled launches of the fused kernel since the last reset: their summed
 * device time (ms, HIP events around each launch), their count and the
 * world-steps they took.  No counterpart in the reference (measurement). */
int r2d_ctx_fused_times(r2d_ctx* ctx, double* kernel_ms, uint64_t* launches, uint64_t* world_steps);
/* The same launches' time per phase of a world's step, summed over the
 * worlds' steps (ms; wall clock of each workgroup): out_ms[8] = cell ranges,
 * buckets, raw candidates, narrowphase, contact points, schedule, solve, 0. */
int r2d_ctx_fused_phase_times(r2d_ctx* ctx, double* out_ms);
/* The same launches' work, summed over their world-steps: out[5] = bodies,
 * cell entries, raw candidates, manifolds, solver items (the units of the
 * step's algorithmic bytes). */
int r2d_ctx_fused_counts(r2d_ctx* ctx, uint64_t* out);
/* Whole steps per workgroup (default on): a batch whose worlds all have at
 * most 1,024 bodies, the reference order and hash tables that fit LDS is
 * stepped by one kernel per call -- every step of every world of the call,
 * one workgroup per world -- instead of the phase kernels of each step.
 * Results are bit-identical either way; off forces the phase kernels.  The
 * environment variable R2D_NO_FUSED=1 sets the default to off.  No
 * counterpart in the reference (Solver.process is one world, one step). */
int r2d_ctx_set_fused(r2d_ctx* ctx, int on);
/* Launch configuration of the solve kernel: -1 (default) is variant 1;
 * 0..12 select the register-slot kernel's slots x workgroup size, the record
 * kernel's workgroup size, and the fused kernel's shape (table in
 * r2d_kernels.hip; 9-12 solve with records only).
 * Every variant gives the same bits; the autotuner (resolve2d_amd/autotune.py)
 * picks the fastest for a batch shape. */
int r2d_ctx_set_solve_variant(r2d_ctx* ctx, int variant);
/* Largest world (bodies) whose solver body state is kept in LDS; larger
 * worlds use the solve kernel with body state in global memory.  0 (default)
 * = as many as LDS holds.  Results are identical either way; this is a
 * launch-shape knob for the autotuner. */
int r2d_ctx_set_lds_limit(r2d_ctx* ctx, uint32_t max_bodies);
/* Phase kernels only: a batch of at most 64 worlds whose solver body state
 * fits LDS takes the record kernel with 512 lanes per world instead of the
 * variant's shape (on, the default: such worlds are alone on their CUs and
 * latency-bound).  Off keeps the variant's record kernel.  Same bits either
 * way; a launch-shape knob with no counterpart in the reference. */
int r2d_ctx_set_wide_record(r2d_ctx* ctx, int on);
/* Totals of the last step of the context: entries, raw candidates,
 * manifolds, solver items (out[4]). */
int r2d_ctx_last_counts(r2d_ctx* ctx, uint64_t* out);

/* Text of the last failure on the calling thread (which HIP call, which
 * world), for R2D_E_DEVICE and friends. */
int r2d_last_error(char* buf, size_t n);

/* Test hook with no counterpart in the reference: evaluates the device's
 * scalar arithmetic on n inputs so tests can compare its bits with the CPU
 * restatement.  op 0 sin, 1 cos (the reference's rotate2 trig,
 * src/core/nmath.zig:129-136), 2 sqrt, 3 reciprocal, 4 division of input
 * pairs (in[2i] / in[2i+1]; `in` holds 2n floats). */
int r2d_debug_eval(r2d_ctx* ctx, int op, const float* in, float* out, size_t n);

/* Measurement hook with no counterpart in the reference: the device's
 * measured peaks, for rooflines against what the hardware delivers.
 * out[0]: HBM read stream (GB/s, 1 GiB read by 16-byte loads, 4 in flight
 * per thread); out[1]: HBM copy (GB/s of read + written bytes, 1 GiB);
 * out[2]: LDS reads (GB/s over all CUs, conflict-free 16-byte reads).
 * Allocates and frees 2 GiB of device memory and waits for the device. */
int r2d_debug_peaks(r2d_ctx* ctx, double* out);

/* ------------------------------------------------------- bulk ingestion */
/* n bodies in one call, in order (ids as n calls of r2d_make_disc /
 * r2d_make_rect would assign them, EntityFactory.appendBody lib.zig:66-71):
 * kind[i] 0 = disc (dim_a = radius), 1 = rectangle (dim_a x dim_b).
 * ids_out may be NULL.  From 32 bodies on, the device builds them: the
 * inputs are copied to the device and one kernel writes mass, inertia,
 * momenta and AABBs into the world's resident state (the same bits as the
 * single calls); the call returns once the inputs have been read. */
int r2d_make_bodies(r2d_world* w, size_t n, const r2d_body_options* bo, const int32_t* kind, const float* dim_a,
                    const float* dim_b, uint32_t* ids_out);

/* ------------------------------------------------------------- serving */
/* No counterpart in the reference (it steps one world per call).  A server
 * steps the worlds of a context's callers with continuous batching:
 * requests -- a built world and a number of steps -- are admitted while the
 * sum of their footprints (r2d_world_footprint) fits hbm_budget_bytes and
 * fewer than max_worlds run.  Admitted worlds are stepped in lanes by size
 * (split_bodies), each lane with its own stream and step scratch, so small
 * worlds are not held to the step time of the largest; within a lane,
 * requests are admitted in arrival order.  Each r2d_serve_tick collects the
 * lanes whose batch step has finished, retires finished requests, admits
 * waiting ones and launches one step of every running world of each idle
 * lane as one batch.  A world's state stays on the device from admission to
 * completion; while a request holds it, edits of the world return
 * R2D_E_BUSY (reads are allowed).  Each world's result is bit-identical to
 * stepping it alone. */
typedef struct r2d_server r2d_server;

enum { R2D_REQ_QUEUED = 0, R2D_REQ_RUNNING = 1, R2D_REQ_DONE = 2, R2D_REQ_FAILED = 3 };

typedef struct {
    size_t hbm_budget_bytes;    /* admission budget over resident footprints */
    uint32_t max_worlds;        /* batch slots */
    float dt;                   /* Solver.process arguments of every step */
    uint32_t sub_steps;
    uint32_t collision_iters;
    uint32_t split_bodies[2];   /* lanes by world size: n <= split[0], n <= split[1], larger;
                                   {0, 0} = one lane.  Each lane has its own stream and scratch
                                   and steps at its own pace. */
    uint32_t steps_per_tick;    /* steps a lane's launch takes for its running worlds: the
                                   smaller of this and the fewest steps any of them has left
                                   (0 or 1: one step per launch).  More steps per launch
                                   amortise launch and collection; admissions wait for the
                                   launch in flight. */
} r2d_serve_options;

typedef struct {
    uint32_t admitted;          /* requests admitted by this tick */
    uint32_t completed;         /* requests finished by the step it collected */
    uint32_t queued;            /* waiting for admission */
    uint32_t running;           /* admitted and not finished */
    uint32_t launched_worlds;   /* worlds in the batch step this tick launched */
    uint64_t launched_bodies;
    size_t resident_bytes;
    double step_gpu_ms;         /* device time of the slowest batch step it collected */
    double step_wall_ms;        /* launch-to-collection wall time of that step */
} r2d_serve_tick_info;

typedef struct {
    int32_t state;              /* R2D_REQ_* */
    int32_t status;             /* r2d_status of a failed request */
    uint32_t steps_done;
    uint32_t steps_requested;
    size_t footprint_bytes;
    uint32_t lane;
    double wait_ms;             /* submit -> admission (so far, while queued) */
    double total_ms;            /* submit -> completion (so far, while not done) */
    double max_step_ms;         /* slowest device step it took part in */
} r2d_request_status;

typedef struct {
    uint64_t submitted, completed, rejected, queued, running;
    uint64_t batch_steps;       /* batched launches collected */
    uint64_t scene_steps;       /* world-steps over them */
    uint64_t body_steps;
    double gpu_ms;              /* device time of all batch steps */
    double wall_ms;             /* launch-to-collection wall time of all batch steps */
    double p50_step_ms, p99_step_ms;    /* device step time seen by each world-step */
    double p50_batch_ms, p99_batch_ms;  /* device time per batch step */
    size_t resident_bytes, peak_resident_bytes;
    uint32_t lanes;
} r2d_serve_summary;

/* Bytes a world is charged at admission: its device state plus its share of
 * the step scratch at the library's initial per-body estimates. */
int r2d_world_footprint(r2d_world* w, size_t* bytes);
int r2d_serve_create(r2d_ctx* ctx, const r2d_serve_options* opt, r2d_server** out);
/* Waits for the launched step; worlds of unfinished requests are released. */
int r2d_serve_destroy(r2d_server* sv);
/* Queues a request; R2D_E_CAPACITY if the world can never fit the budget,
 * R2D_E_INVALID_ID if a constraint names a removed body. */
int r2d_serve_submit(r2d_server* sv, r2d_world* w, uint32_t steps, uint64_t* ticket);
/* Waits until some lane's step has finished (if any is in flight), then
 * admits and launches; it does not wait for the steps it launches, so host
 * work between ticks overlaps the device.  info may be NULL. */
int r2d_serve_tick(r2d_server* sv, r2d_serve_tick_info* info);
/* Ticks until every request has finished. */
int r2d_serve_drain(r2d_server* sv);
int r2d_serve_poll(r2d_server* sv, uint64_t ticket, r2d_request_status* out);
/* Tickets of the requests that finished (done or failed) since the last call,
 * oldest first: up to cap of them into tickets, their number in *n; the rest
 * wait for the next call.  Lets a serving loop visit only finished requests
 * instead of polling every one it holds. */
int r2d_serve_completed(r2d_server* sv, uint64_t* tickets, uint32_t cap, uint32_t* n);
int r2d_serve_stats(r2d_server* sv, r2d_serve_summary* out);

#ifdef __cplusplus
}
#endif

#endif /* R2D_H */
