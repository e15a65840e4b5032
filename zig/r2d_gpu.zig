//! Zig binding of the MI355X step library (libr2d_hip.so, C ABI in include/r2d.h).
//!
//! A drop-in for resolve2d's `Solver` / `EntityFactory` surface
//! (src/root.zig:4-10, src/core/lib.zig:13-315): the same names, argument
//! meaning and errors, with the step (`Solver.process`, lib.zig:189-251) run
//! on the GPU.  Place this file at src/gpu/r2d_gpu.zig in the resolve2d tree
//! (it takes the reference's own Vector2 and Constraint.Parameters), and in
//! build.zig:
//!
//!     mod.addIncludePath(b.path("path/to/include"));
//!     mod.addLibraryPath(b.path("path/to/resolve2d_amd/lib"));
//!     mod.linkSystemLibrary("r2d_hip", .{});
//!     mod.link_libc = true;
//!
//! Differences a caller sees, each forced by the device boundary:
//!   - `BodyHandle.body()` returns a copy of the body (`?RigidBody`) read from
//!     the device, not a pointer into the solver's map; after editing
//!     `props` or `static`, call `commit()` to write it back (the reference's
//!     callers write through the pointer: demos/native/src/main.zig:114-135).
//!   - ids are u32 (RigidBody.Id is u16, RigidBody.zig:15), so worlds may
//!     exceed 65,535 bodies; ids are still never reused.
//!   - `process` checks joints before stepping: a joint naming a removed body
//!     returns error.InvalidRigidBodyId with the world unstepped (the
//!     reference raises it from inside the constraint loop).
//!   - `clear` also empties the excluded pairs (the reference's clear reuses a
//!     deinit'ed map, lib.zig:181-187); the id counter keeps counting.
//!   - the allocator arguments of `init`, `process` and `clear` are accepted
//!     for signature compatibility; the library owns its memory (host mirror
//!     and device state).
//!
//! Beyond the reference's surface: bulk ingestion (`makeBodies`), readback
//! (`Solver.getBodies` / `getManifolds`), the tuned hash, warm starting,
//! the coloured schedule, the step without allocation (`processStatic`) and
//! continuous batching (`Server`).
//!
//! No Zig toolchain exists where this library is built, so this file is
//! written against include/r2d.h and has not been compiled there.
const std = @import("std");
const Allocator = std.mem.Allocator;
const nmath = @import("../core/nmath.zig");
const Vector2 = nmath.Vector2;
const Constraint = @import("../core/Constraints/Constraint.zig");

// ------------------------------------------------------------------ C ABI
pub const Ctx = opaque {};
pub const World = opaque {};

const CBodyOptions = extern struct {
    pos_x: f32,
    pos_y: f32,
    vel_x: f32,
    vel_y: f32,
    angle: f32,
    omega: f32,
    mu: f32,
    mass_kind: i32, // 0: density (kg/m^2), 1: mass (kg)
    mass_value: f32,
};

const CParams = extern struct {
    power_max: f32,
    power_min: f32,
    beta: f32,
};

pub const CBodyState = extern struct {
    id: u32,
    type: u32, // 0 disc, 1 rectangle
    is_static: u32,
    num_normals: u32,
    pos_x: f32,
    pos_y: f32,
    mom_x: f32,
    mom_y: f32,
    force_x: f32,
    force_y: f32,
    mass: f32,
    angle: f32,
    ang_momentum: f32,
    torque: f32,
    inertia: f32,
    mu: f32,
    aabb_x: f32,
    aabb_y: f32,
    aabb_half_width: f32,
    aabb_half_height: f32,
    radius: f32,
    width: f32,
    height: f32,
};

extern fn r2d_ctx_create(device: c_int, out: *?*Ctx) c_int;
extern fn r2d_ctx_destroy(ctx: *Ctx) c_int;
extern fn r2d_world_create(ctx: *Ctx, cell_width: f32, table_mult: usize, out: *?*World) c_int;
extern fn r2d_world_destroy(w: *World) c_int;
extern fn r2d_world_clear(w: *World) c_int;
extern fn r2d_make_disc(w: *World, bo: *const CBodyOptions, radius: f32, id: *u32) c_int;
extern fn r2d_make_rect(w: *World, bo: *const CBodyOptions, width: f32, height: f32, id: *u32) c_int;
extern fn r2d_set_static(w: *World, id: u32, is_static: c_int) c_int;
extern fn r2d_remove_body(w: *World, id: u32) c_int;
extern fn r2d_make_gravity(w: *World, g: f32) c_int;
extern fn r2d_make_spring(w: *World, id: u32, k: f32, c: f32, ax: f32, ay: f32) c_int;
extern fn r2d_make_distance_joint(w: *World, p: *const CParams, id1: u32, id2: u32, d: f32, h: *usize) c_int;
extern fn r2d_make_offset_distance_joint(w: *World, p: *const CParams, id1: u32, id2: u32, r1x: f32, r1y: f32, r2x: f32, r2y: f32, d: f32, h: *usize) c_int;
extern fn r2d_make_fixed_position_joint(w: *World, p: *const CParams, id: u32, tx: f32, ty: f32, h: *usize) c_int;
extern fn r2d_make_motor_joint(w: *World, p: *const CParams, id: u32, omega: f32, h: *usize) c_int;
extern fn r2d_exclude_collision_pair(w: *World, id1: u32, id2: u32) c_int;
extern fn r2d_process(w: *World, dt: f32, sub_steps: usize, collision_iters: usize) c_int;
extern fn r2d_process_batch(ctx: *Ctx, worlds: [*]const *World, n: usize, dt: f32, sub_steps: usize, collision_iters: usize, steps: usize) c_int;
extern fn r2d_num_bodies(w: *World, n: *usize) c_int;
extern fn r2d_body_id_at(w: *World, i: usize, id: *u32) c_int;
extern fn r2d_get_body(w: *World, id: u32, out: *CBodyState) c_int;
extern fn r2d_set_body(w: *World, id: u32, in: *const CBodyState) c_int;
extern fn r2d_set_warm_start(w: *World, factor: f32) c_int;
extern fn r2d_set_schedule(w: *World, mode: c_int) c_int;
extern fn r2d_set_hash_params(w: *World, mode: c_int, cell_width: f32, table_mult: usize) c_int;
extern fn r2d_process_static(w: *World, dt: f32, sub_steps: usize, collision_iters: usize) c_int;
extern fn r2d_make_bodies(w: *World, n: usize, bo: [*]const CBodyOptions, kind: [*]const i32, dim_a: [*]const f32, dim_b: [*]const f32, ids_out: ?[*]u32) c_int;
extern fn r2d_get_bodies(w: *World, out: [*]CBodyState, cap: usize, n_out: *usize) c_int;
extern fn r2d_get_manifolds(w: *World, out: [*]CManifoldState, cap: usize, n_out: *usize) c_int;
extern fn r2d_last_error(buf: [*]u8, n: usize) c_int;
extern fn r2d_world_footprint(w: *World, bytes: *usize) c_int;
extern fn r2d_ctx_sync(ctx: *Ctx) c_int;
extern fn r2d_ctx_set_solve_variant(ctx: *Ctx, variant: c_int) c_int;
extern fn r2d_ctx_set_fused(ctx: *Ctx, on: c_int) c_int;
extern fn r2d_ctx_set_profiling(ctx: *Ctx, on: c_int) c_int;
extern fn r2d_ctx_reset_phase_times(ctx: *Ctx) c_int;
extern fn r2d_ctx_fused_times(ctx: *Ctx, kernel_ms: *f64, launches: *u64, world_steps: *u64) c_int;
extern fn r2d_ctx_step_times(ctx: *Ctx, out: [*]f32, cap: usize, n: *usize) c_int;
extern fn r2d_debug_peaks(ctx: *Ctx, out: *[3]f64) c_int;
extern fn r2d_serve_create(ctx: *Ctx, opt: *const ServeOptions, out: *?*Server.Handle) c_int;
extern fn r2d_serve_destroy(sv: *Server.Handle) c_int;
extern fn r2d_serve_submit(sv: *Server.Handle, w: *World, steps: u32, ticket: *u64) c_int;
extern fn r2d_serve_tick(sv: *Server.Handle, info: ?*ServeTickInfo) c_int;
extern fn r2d_serve_drain(sv: *Server.Handle) c_int;
extern fn r2d_serve_poll(sv: *Server.Handle, ticket: u64, out: *RequestStatus) c_int;
extern fn r2d_serve_completed(sv: *Server.Handle, tickets: [*]u64, cap: u32, n: *u32) c_int;
extern fn r2d_serve_stats(sv: *Server.Handle, out: *ServeSummary) c_int;

/// One contact point of a manifold (collision.zig:22-51), as r2d_get_manifolds reports it.
pub const CManifoldPoint = extern struct {
    valid: u32,
    acc_pn: f32,
    acc_pt: f32,
    ref_r_x: f32,
    ref_r_y: f32,
    inc_r_x: f32,
    inc_r_y: f32,
    pos_x: f32,
    pos_y: f32,
    depth: f32,
    original_depth: f32,
    mass_n: f32,
    mass_t: f32,
};

/// One contact manifold (collision.zig:54-71).
pub const CManifoldState = extern struct {
    ref_id: u32,
    inc_id: u32,
    normal_x: f32,
    normal_y: f32,
    tangent_x: f32,
    tangent_y: f32,
    friction: f32,
    points: [2]CManifoldPoint,
};

/// r2d_serve_options (include/r2d.h).
pub const ServeOptions = extern struct {
    hbm_budget_bytes: usize,
    max_worlds: u32,
    dt: f32,
    sub_steps: u32,
    collision_iters: u32,
    split_bodies: [2]u32,
    steps_per_tick: u32,
};

pub const ServeTickInfo = extern struct {
    admitted: u32,
    completed: u32,
    queued: u32,
    running: u32,
    launched_worlds: u32,
    launched_bodies: u64,
    resident_bytes: usize,
    step_gpu_ms: f64,
    step_wall_ms: f64,
};

pub const RequestState = enum(i32) { queued = 0, running = 1, done = 2, failed = 3 };

pub const RequestStatus = extern struct {
    state: i32, // RequestState
    status: i32,
    steps_done: u32,
    steps_requested: u32,
    footprint_bytes: usize,
    lane: u32,
    wait_ms: f64,
    total_ms: f64,
    max_step_ms: f64,
};

pub const ServeSummary = extern struct {
    submitted: u64,
    completed: u64,
    rejected: u64,
    queued: u64,
    running: u64,
    batch_steps: u64,
    scene_steps: u64,
    body_steps: u64,
    gpu_ms: f64,
    wall_ms: f64,
    p50_step_ms: f64,
    p99_step_ms: f64,
    p50_batch_ms: f64,
    p99_batch_ms: f64,
    resident_bytes: usize,
    peak_resident_bytes: usize,
    lanes: u32,
};

/// The text of the calling thread's last failure (which HIP call, which world).
pub fn lastError(buf: []u8) []const u8 {
    if (buf.len == 0 or r2d_last_error(buf.ptr, buf.len) != 0) return buf[0..0];
    return std.mem.sliceTo(buf, 0);
}

pub const Error = error{ InvalidRigidBodyId, NoSuchIdExists, OutOfMemory, NonFinite, Device, Busy, InvalidArgument, OutOfRange, Capacity };

fn check(rc: c_int) Error!void {
    return switch (rc) {
        0 => {},
        1 => error.InvalidArgument,
        2 => error.InvalidRigidBodyId,
        3 => error.OutOfMemory,
        4 => error.NonFinite,
        7 => error.OutOfRange,
        8 => error.Capacity,
        9 => error.Busy,
        else => error.Device,
    };
}

fn params(p: Constraint.Parameters) CParams {
    return .{ .power_max = p.power_max, .power_min = p.power_min, .beta = p.beta };
}

// ----------------------------------------------------- one device context
/// A GPU and its stream.  Worlds (Solvers) of one context can be stepped
/// together in one launch with `processBatch`.
pub const Device = struct {
    ctx: *Ctx,

    pub fn open(device: c_int) Error!Device {
        var c: ?*Ctx = null;
        try check(r2d_ctx_create(device, &c));
        return .{ .ctx = c.? };
    }

    /// Destroy every Solver of the device first (error.Busy otherwise).
    pub fn close(self: *Device) Error!void {
        try check(r2d_ctx_destroy(self.ctx));
    }

    /// `steps` steps of every solver: the same result as calling each one's
    /// process `steps` times, in one launch for worlds of up to 1,024 bodies.
    pub fn processBatch(self: *Device, solvers: []const *Solver, dt: f32, sub_steps: usize, collision_iters: usize, steps: usize) Error!void {
        var buf: [4096]*World = undefined;
        var done: usize = 0;
        while (done < solvers.len) {
            const n = @min(buf.len, solvers.len - done);
            for (solvers[done .. done + n], 0..) |s, i| buf[i] = s.world;
            try check(r2d_process_batch(self.ctx, &buf, n, dt, sub_steps, collision_iters, steps));
            done += n;
        }
    }

    /// Waits for all device work of the context.
    pub fn sync(self: *Device) Error!void {
        try check(r2d_ctx_sync(self.ctx));
    }

    /// Launch-shape knobs (results are the same bits for every setting):
    /// the solve variant (-1: the library's default), whole steps per
    /// workgroup on or off.
    pub fn setSolveVariant(self: *Device, variant: c_int) Error!void {
        try check(r2d_ctx_set_solve_variant(self.ctx, variant));
    }

    pub fn setFused(self: *Device, on: bool) Error!void {
        try check(r2d_ctx_set_fused(self.ctx, @intFromBool(on)));
    }

    /// Device timing of the following calls (per step and per fused launch).
    pub fn setProfiling(self: *Device, on: bool) Error!void {
        try check(r2d_ctx_set_profiling(self.ctx, @intFromBool(on)));
    }

    pub fn resetPhaseTimes(self: *Device) Error!void {
        try check(r2d_ctx_reset_phase_times(self.ctx));
    }

    pub const FusedTimes = struct { kernel_ms: f64, launches: u64, world_steps: u64 };

    /// Device time of the profiled fused launches since the last reset.
    pub fn fusedTimes(self: *Device) Error!FusedTimes {
        var t: FusedTimes = .{ .kernel_ms = 0, .launches = 0, .world_steps = 0 };
        try check(r2d_ctx_fused_times(self.ctx, &t.kernel_ms, &t.launches, &t.world_steps));
        return t;
    }

    /// Device time (ms) of every profiled step since the last reset, into `buf`.
    pub fn stepTimes(self: *Device, buf: []f32) Error![]f32 {
        var n: usize = 0;
        try check(r2d_ctx_step_times(self.ctx, buf.ptr, buf.len, &n));
        return buf[0..@min(n, buf.len)];
    }

    pub const Peaks = struct { hbm_read: f64, hbm_copy: f64, lds_read: f64 };

    /// The device's measured HBM read / copy and LDS read rates (GB/s).
    pub fn measuredPeaks(self: *Device) Error!Peaks {
        var out: [3]f64 = undefined;
        try check(r2d_debug_peaks(self.ctx, &out));
        return .{ .hbm_read = out[0], .hbm_copy = out[1], .lds_read = out[2] };
    }
};

// ------------------------------------------------------------------ bodies
/// RigidBody.Props (RigidBody.zig:69-81).
pub const Props = struct {
    pos: Vector2,
    momentum: Vector2,
    force: Vector2,
    mass: f32,
    angle: f32,
    ang_momentum: f32,
    torque: f32,
    inertia: f32,
    mu: f32,
};

pub const Type = enum { disc, rectangle };

/// A copy of one body as the device holds it (RigidBody.zig:83-91).
pub const RigidBody = struct {
    solver: *Solver,
    id: u32,
    static: bool,
    type: Type,
    num_normals: usize,
    props: Props,
    aabb_pos: Vector2,
    aabb_half_width: f32,
    aabb_half_height: f32,
    radius: f32, // discs
    width: f32, // rectangles
    height: f32,

    fn fromState(solver: *Solver, s: CBodyState) RigidBody {
        return .{
            .solver = solver,
            .id = s.id,
            .static = s.is_static != 0,
            .type = if (s.type == 0) .disc else .rectangle,
            .num_normals = s.num_normals,
            .props = .{
                .pos = Vector2.init(s.pos_x, s.pos_y),
                .momentum = Vector2.init(s.mom_x, s.mom_y),
                .force = Vector2.init(s.force_x, s.force_y),
                .mass = s.mass,
                .angle = s.angle,
                .ang_momentum = s.ang_momentum,
                .torque = s.torque,
                .inertia = s.inertia,
                .mu = s.mu,
            },
            .aabb_pos = Vector2.init(s.aabb_x, s.aabb_y),
            .aabb_half_width = s.aabb_half_width,
            .aabb_half_height = s.aabb_half_height,
            .radius = s.radius,
            .width = s.width,
            .height = s.height,
        };
    }

    /// Writes `static` and the writable props back: position, momentum,
    /// force, angle, angular momentum, torque (the setters of
    /// src/wasm_root.zig:156-224 plus position and angle).
    pub fn commit(self: *const RigidBody) Error!void {
        var s: CBodyState = undefined;
        try check(r2d_get_body(self.solver.world, self.id, &s));
        s.pos_x = self.props.pos.x;
        s.pos_y = self.props.pos.y;
        s.mom_x = self.props.momentum.x;
        s.mom_y = self.props.momentum.y;
        s.force_x = self.props.force.x;
        s.force_y = self.props.force.y;
        s.angle = self.props.angle;
        s.ang_momentum = self.props.ang_momentum;
        s.torque = self.props.torque;
        try check(r2d_set_body(self.solver.world, self.id, &s));
        if ((s.is_static != 0) != self.static)
            try check(r2d_set_static(self.solver.world, self.id, @intFromBool(self.static)));
    }
};

// ---------------------------------------------------------- EntityFactory
/// EntityFactory (lib.zig:13-130).
pub const EntityFactory = struct {
    /// EntityFactory.BodyHandle (lib.zig:15-32).
    pub const BodyHandle = struct {
        id: u32,
        solver: *Solver,

        /// The body as the device holds it, or null when the id is gone.
        pub fn body(self: BodyHandle) ?RigidBody {
            var s: CBodyState = undefined;
            if (r2d_get_body(self.solver.world, self.id, &s) != 0) return null;
            return RigidBody.fromState(self.solver, s);
        }

        pub fn body_unwrap(self: BodyHandle) RigidBody {
            return self.body() orelse @panic("Tried to unwrap a null *RigidBody");
        }
    };
    pub const ConstraintHandle = usize;

    /// EntityFactory.BodyOptions (lib.zig:35-46).
    pub const BodyOptions = struct {
        pos: Vector2,
        vel: Vector2 = .{},
        angle: f32 = 0.0,
        omega: f32 = 0.0,
        mu: f32 = 0.5,
        mass_prop: union(enum) {
            density: f32,
            mass: f32,
        },

        fn toC(bo: BodyOptions) CBodyOptions {
            return .{
                .pos_x = bo.pos.x,
                .pos_y = bo.pos.y,
                .vel_x = bo.vel.x,
                .vel_y = bo.vel.y,
                .angle = bo.angle,
                .omega = bo.omega,
                .mu = bo.mu,
                .mass_kind = switch (bo.mass_prop) {
                    .density => 0,
                    .mass => 1,
                },
                .mass_value = switch (bo.mass_prop) {
                    .density => |d| d,
                    .mass => |m| m,
                },
            };
        }
    };

    pub const DiscOptions = struct { radius: f32 = 1.0 };
    pub const RectangleOptions = struct { width: f32 = 1.0, height: f32 = 0.5 };

    solver: *Solver,

    const Self = @This();

    pub fn makeDiscBody(self: *Self, bo: BodyOptions, go: DiscOptions) Error!BodyHandle {
        const c = bo.toC();
        var id: u32 = 0;
        try check(r2d_make_disc(self.solver.world, &c, go.radius, &id));
        return .{ .id = id, .solver = self.solver };
    }

    pub fn makeRectangleBody(self: *Self, bo: BodyOptions, go: RectangleOptions) Error!BodyHandle {
        const c = bo.toC();
        var id: u32 = 0;
        try check(r2d_make_rect(self.solver.world, &c, go.width, go.height, &id));
        return .{ .id = id, .solver = self.solver };
    }

    pub fn makeDownwardsGravity(self: *Self, g: f32) Error!void {
        try check(r2d_make_gravity(self.solver.world, g));
    }

    /// Anchored spring F += -k (x - anchor) - c v on one body.  Not in the
    /// reference (its force generators hold only gravity, ForceGen.zig:12-14).
    pub fn makeSpring(self: *Self, h: BodyHandle, k: f32, c: f32, anchor: Vector2) Error!void {
        try check(r2d_make_spring(self.solver.world, h.id, k, c, anchor.x, anchor.y));
    }

    pub fn makeOffsetDistanceJoint(self: *Self, p: Constraint.Parameters, h1: BodyHandle, h2: BodyHandle, r1: Vector2, r2: Vector2, target_distance: f32) Error!ConstraintHandle {
        const c = params(p);
        var out: usize = 0;
        try check(r2d_make_offset_distance_joint(self.solver.world, &c, h1.id, h2.id, r1.x, r1.y, r2.x, r2.y, target_distance, &out));
        return out;
    }

    pub fn makeDistanceJoint(self: *Self, p: Constraint.Parameters, h1: BodyHandle, h2: BodyHandle, target_distance: f32) Error!ConstraintHandle {
        const c = params(p);
        var out: usize = 0;
        try check(r2d_make_distance_joint(self.solver.world, &c, h1.id, h2.id, target_distance, &out));
        return out;
    }

    pub fn makeFixedPositionJoint(self: *Self, p: Constraint.Parameters, h: BodyHandle, target_position: Vector2) Error!ConstraintHandle {
        const c = params(p);
        var out: usize = 0;
        try check(r2d_make_fixed_position_joint(self.solver.world, &c, h.id, target_position.x, target_position.y, &out));
        return out;
    }

    pub fn makeMotorJoint(self: *Self, p: Constraint.Parameters, h: BodyHandle, target_omega: f32) Error!ConstraintHandle {
        const c = params(p);
        var out: usize = 0;
        try check(r2d_make_motor_joint(self.solver.world, &c, h.id, target_omega, &out));
        return out;
    }

    pub fn excludeCollisionPair(self: *Self, h1: BodyHandle, h2: BodyHandle) Error!void {
        try check(r2d_exclude_collision_pair(self.solver.world, h1.id, h2.id));
    }
};

// ------------------------------------------------------------------ Solver
/// Solver (lib.zig:132-315) with its step on the GPU.
pub const Solver = struct {
    world: *World,

    const Self = @This();

    /// Solver.init (lib.zig:145-157) on a device.  Like the reference, the
    /// hash arguments are stored and the step uses cell 4.0 and 2N buckets.
    pub fn init(_: Allocator, device: *Device, spatialhash_cell_width: f32, spatialhash_table_size_mult: usize) Error!Self {
        var w: ?*World = null;
        try check(r2d_world_create(device.ctx, spatialhash_cell_width, spatialhash_table_size_mult, &w));
        return .{ .world = w.? };
    }

    pub fn deinit(self: *Self) void {
        _ = r2d_world_destroy(self.world);
    }

    /// Solver.clear (lib.zig:181-187).
    pub fn clear(self: *Self, _: Allocator) Error!void {
        try check(r2d_world_clear(self.world));
    }

    /// Solver.process (lib.zig:189-251): the GPU step.
    pub fn process(self: *Self, _: Allocator, dt: f32, sub_steps: usize, collision_iters: usize) Error!void {
        try check(r2d_process(self.world, dt, sub_steps, collision_iters));
    }

    /// Solver.bodyHandle (lib.zig:301-306).
    pub fn bodyHandle(self: *Self, id: u32) EntityFactory.BodyHandle {
        return .{ .id = id, .solver = self };
    }

    /// Solver.removeRigidBody (lib.zig:308-310): swapRemove order.
    pub fn removeRigidBody(self: *Self, id: u32) Error!void {
        const rc = r2d_remove_body(self.world, id);
        if (rc == 2) return error.NoSuchIdExists;
        try check(rc);
    }

    /// Solver.entityFactory (lib.zig:312-314).
    pub fn entityFactory(self: *Self) EntityFactory {
        return .{ .solver = self };
    }

    /// solverGetNumBodies (src/wasm_root.zig:51-53).
    pub fn numBodies(self: *Self) Error!usize {
        var n: usize = 0;
        try check(r2d_num_bodies(self.world, &n));
        return n;
    }

    /// solverGetBodyIdBasedOnIter (src/wasm_root.zig:55-58): iteration order.
    pub fn bodyIdAt(self: *Self, i: usize) Error!u32 {
        var id: u32 = 0;
        try check(r2d_body_id_at(self.world, i, &id));
        return id;
    }

    /// Opt-in warm starting (0 keeps the reference's results).
    pub fn setWarmStart(self: *Self, factor: f32) Error!void {
        try check(r2d_set_warm_start(self.world, factor));
    }

    /// 0: the reference's order (default); 1: graph-coloured batches.
    pub fn setSchedule(self: *Self, mode: c_int) Error!void {
        try check(r2d_set_schedule(self.world, mode));
    }

    /// Spatial-hash parameters: mode 0 the reference's cell 4.0 and 2N
    /// buckets (lib.zig:254-255); mode 1 the given cell width and
    /// table_mult * N buckets (README.md:62-63).
    pub fn setHashParams(self: *Self, mode: c_int, cell_width: f32, table_mult: usize) Error!void {
        try check(r2d_set_hash_params(self.world, mode, cell_width, table_mult));
    }

    /// `process` for callers that cannot allocate device memory: the step's
    /// scratch is a fixed device array of the library (worlds of at most
    /// 1,024 bodies, reference order, no warm starting; error.Capacity
    /// beyond).  Calls on one GPU must not overlap.
    pub fn processStatic(self: *Self, dt: f32, sub_steps: usize, collision_iters: usize) Error!void {
        try check(r2d_process_static(self.world, dt, sub_steps, collision_iters));
    }

    /// All bodies in iteration order into `buf`; the slice of it filled, or
    /// error.OutOfRange when `buf` is too short (numBodies gives the size).
    pub fn getBodies(self: *Self, buf: []CBodyState) Error![]CBodyState {
        var n: usize = 0;
        try check(r2d_get_bodies(self.world, buf.ptr, buf.len, &n));
        if (n > buf.len) return error.OutOfRange;
        return buf[0..n];
    }

    /// The last step's manifolds in the reference's order (Solver.manifolds);
    /// error.OutOfRange when `buf` is too short.
    pub fn getManifolds(self: *Self, buf: []CManifoldState) Error![]CManifoldState {
        var n: usize = 0;
        try check(r2d_get_manifolds(self.world, buf.ptr, buf.len, &n));
        if (n > buf.len) return error.OutOfRange;
        return buf[0..n];
    }

    /// Device bytes a server charges this world at admission.
    pub fn footprint(self: *Self) Error!usize {
        var b: usize = 0;
        try check(r2d_world_footprint(self.world, &b));
        return b;
    }
};

// ----------------------------------------------------------- bulk bodies
/// Many bodies in one call, ids as the same number of makeDiscBody /
/// makeRectangleBody calls would assign them (EntityFactory.appendBody,
/// lib.zig:66-71): kind[i] 0 disc (dim_a = radius), 1 rectangle (dim_a x
/// dim_b).  From 32 bodies on the device builds them.
pub fn makeBodies(solver: *Solver, bo: []const EntityFactory.BodyOptions, kind: []const i32, dim_a: []const f32, dim_b: []const f32, ids_out: ?[]u32, scratch: Allocator) Error!void {
    const n = bo.len;
    if (kind.len != n or dim_a.len != n or dim_b.len != n) return error.InvalidArgument;
    if (ids_out) |ids| if (ids.len != n) return error.InvalidArgument;
    const cbo = scratch.alloc(CBodyOptions, n) catch return error.OutOfMemory;
    defer scratch.free(cbo);
    for (bo, cbo) |b, *c| c.* = b.toC();
    try check(r2d_make_bodies(solver.world, n, cbo.ptr, kind.ptr, dim_a.ptr, dim_b.ptr, if (ids_out) |ids| ids.ptr else null));
}

// ----------------------------------------------------------------- serving
/// Continuous batching of many worlds on one device (no counterpart in the
/// reference): requests -- a built Solver and a number of steps -- are
/// admitted while their footprints fit the HBM budget, stepped in lanes by
/// world size, and each world's result is bit-identical to stepping it
/// alone.  The solver of a running request refuses edits (error.Busy).
pub const Server = struct {
    pub const Handle = opaque {};
    handle: *Handle,

    pub fn create(device: *Device, opt: ServeOptions) Error!Server {
        var h: ?*Handle = null;
        try check(r2d_serve_create(device.ctx, &opt, &h));
        return .{ .handle = h.? };
    }

    pub fn destroy(self: *Server) Error!void {
        try check(r2d_serve_destroy(self.handle));
    }

    /// Queues `steps` steps of the solver; returns its ticket.
    pub fn submit(self: *Server, solver: *Solver, steps: u32) Error!u64 {
        var t: u64 = 0;
        try check(r2d_serve_submit(self.handle, solver.world, steps, &t));
        return t;
    }

    /// Collects finished lane steps, retires, admits and launches.
    pub fn tick(self: *Server) Error!ServeTickInfo {
        var info: ServeTickInfo = undefined;
        try check(r2d_serve_tick(self.handle, &info));
        return info;
    }

    pub fn drain(self: *Server) Error!void {
        try check(r2d_serve_drain(self.handle));
    }

    pub fn poll(self: *Server, ticket: u64) Error!RequestStatus {
        var st: RequestStatus = undefined;
        try check(r2d_serve_poll(self.handle, ticket, &st));
        return st;
    }

    /// Tickets finished since the last call, oldest first, into `buf`.
    pub fn completed(self: *Server, buf: []u64) Error![]u64 {
        var n: u32 = 0;
        try check(r2d_serve_completed(self.handle, buf.ptr, @intCast(buf.len), &n));
        return buf[0..n];
    }

    pub fn stats(self: *Server) Error!ServeSummary {
        var s: ServeSummary = undefined;
        try check(r2d_serve_stats(self.handle, &s));
        return s;
    }
};

test "a box rests on a static ground" {
    var dev = try Device.open(0);
    defer dev.close() catch {};
    var solver = try Solver.init(std.testing.allocator, &dev, 4.0, 2);
    defer solver.deinit();
    var fac = solver.entityFactory();
    try fac.makeDownwardsGravity(9.81);
    const ground = try fac.makeRectangleBody(.{ .pos = Vector2.init(0, -0.5), .mass_prop = .{ .density = 1 } }, .{ .width = 100, .height = 1 });
    var g = ground.body_unwrap();
    g.static = true;
    try g.commit();
    const box = try fac.makeRectangleBody(.{ .pos = Vector2.init(0, 0.5), .mass_prop = .{ .density = 1 } }, .{ .width = 1, .height = 1 });
    for (0..120) |_| try solver.process(std.testing.allocator, 1.0 / 60.0, 4, 4);
    const b = box.body_unwrap();
    try std.testing.expect(@abs(b.props.pos.y - 0.5) < 0.05);
    try solver.removeRigidBody(box.id);
    try std.testing.expect(box.body() == null);
    try std.testing.expectError(error.NoSuchIdExists, solver.removeRigidBody(box.id));
}
