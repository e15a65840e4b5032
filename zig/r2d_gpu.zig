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

pub const Error = error{ InvalidRigidBodyId, NoSuchIdExists, OutOfMemory, NonFinite, Device, Busy, InvalidArgument };

fn check(rc: c_int) Error!void {
    return switch (rc) {
        0 => {},
        1 => error.InvalidArgument,
        2 => error.InvalidRigidBodyId,
        3 => error.OutOfMemory,
        4 => error.NonFinite,
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
