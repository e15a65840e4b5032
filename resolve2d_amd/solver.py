"""Host-side mirror of resolve2d's Zig API (src/root.zig:4-10) over the C ABI.

Names, argument meaning and errors follow the reference:

    Solver.init(cell_width, table_size_mult)          lib.zig:145
    solver.process(dt, sub_steps, collision_iters)    lib.zig:189   (runs on the GPU)
    solver.entityFactory().makeDiscBody(opts, geo)    lib.zig:73
    solver.entityFactory().makeRectangleBody(...)     lib.zig:84
    fac.makeDownwardsGravity(g)                       lib.zig:95
    fac.make{Distance,OffsetDistance,FixedPosition,Motor}Joint(params, ...)  lib.zig:100-122
    fac.excludeCollisionPair(h1, h2)                  lib.zig:124
    solver.bodyHandle(id).body_unwrap()               lib.zig:301, 23-31
    solver.removeRigidBody(id)                        lib.zig:308  (InvalidRigidBodyId if absent)

A BodyHandle's body view reads through to device state and writes back on
attribute assignment (`h.body_unwrap().static = True`, `props.momentum = ...`).
"""
import ctypes
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

from . import _native as N

Vec2 = Tuple[float, float]


class Context:
    """One device + stream; worlds of one context can be stepped together."""

    def __init__(self, device: int = 0):
        self.lib = N.load()
        h = ctypes.c_void_p()
        N.check(self.lib.r2d_ctx_create(device, ctypes.byref(h)), "r2d_ctx_create")
        self.handle = h
        self.profiling = False
        self._servers = {}   # open servers and worlds made on this context, by id()
        self._worlds = {}

    def adopt(self, obj, server: bool = False):
        """Register an open world or server created on this context; close()
        closes what is still open first (servers, then worlds), since the
        library refuses to destroy a context that they still use."""
        (self._servers if server else self._worlds)[id(obj)] = obj

    def release(self, obj):
        """obj has been closed."""
        self._servers.pop(id(obj), None)
        self._worlds.pop(id(obj), None)

    def close(self):
        if self.handle:
            for o in list(self._servers.values()) + list(self._worlds.values()):
                o.close()
            N.check(self.lib.r2d_ctx_destroy(self.handle), "r2d_ctx_destroy")
            self.handle = None

    def sync(self):
        N.check(self.lib.r2d_ctx_sync(self.handle), "r2d_ctx_sync")

    def process_batch(self, solvers: Sequence["Solver"], dt: float, sub_steps: int, collision_iters: int,
                      steps: int = 1):
        arr = (ctypes.c_void_p * len(solvers))(*[s.handle for s in solvers])
        N.check(self.lib.r2d_process_batch(self.handle, arr, len(solvers), dt, sub_steps, collision_iters, steps),
                "r2d_process_batch")

    def process_batch_handles(self, handles, dt: float, sub_steps: int, collision_iters: int, steps: int = 1):
        arr = (ctypes.c_void_p * len(handles))(*handles)
        N.check(self.lib.r2d_process_batch(self.handle, arr, len(handles), dt, sub_steps, collision_iters, steps),
                "r2d_process_batch")

    def prepare_batch(self, handles) -> "PreparedBatch":
        """A reusable handle array for stepping the same worlds many times."""
        return PreparedBatch(self, handles)

    def set_solve_variant(self, variant: int):
        N.check(self.lib.r2d_ctx_set_solve_variant(self.handle, variant), "r2d_ctx_set_solve_variant")

    def set_lds_limit(self, max_bodies: int):
        N.check(self.lib.r2d_ctx_set_lds_limit(self.handle, max_bodies), "r2d_ctx_set_lds_limit")

    def set_wide_record(self, on: bool):
        """Phase kernels: batches of at most 64 LDS-state worlds take the
        512-lane record kernel (default on).  Same bits either way."""
        N.check(self.lib.r2d_ctx_set_wide_record(self.handle, 1 if on else 0), "r2d_ctx_set_wide_record")

    def set_fused(self, on: bool):
        """Whole steps per workgroup for batches of worlds up to 1,024 bodies
        (default on); off forces the phase kernels.  Same bits either way."""
        N.check(self.lib.r2d_ctx_set_fused(self.handle, 1 if on else 0), "r2d_ctx_set_fused")

    def fused_times(self):
        """(kernel ms, launches, world-steps) of profiled fused launches since the last reset."""
        ms = ctypes.c_double()
        launches = ctypes.c_uint64()
        steps = ctypes.c_uint64()
        N.check(self.lib.r2d_ctx_fused_times(self.handle, ctypes.byref(ms), ctypes.byref(launches),
                                             ctypes.byref(steps)), "r2d_ctx_fused_times")
        return ms.value, launches.value, steps.value

    def fused_phase_times(self):
        """Per-phase ms of profiled fused launches, summed over world-steps:
        dict of cell ranges, buckets, candidates, narrowphase, contacts, schedule, solve."""
        out = (ctypes.c_double * 8)()
        N.check(self.lib.r2d_ctx_fused_phase_times(self.handle, out), "r2d_ctx_fused_phase_times")
        names = ["cells", "buckets", "candidates", "narrowphase", "contacts", "schedule", "solve"]
        return {k: out[i] for i, k in enumerate(names)}

    def fused_counts(self):
        """Work of profiled fused launches since the last reset, summed over world-steps."""
        out = (ctypes.c_uint64 * 5)()
        N.check(self.lib.r2d_ctx_fused_counts(self.handle, out), "r2d_ctx_fused_counts")
        return dict(bodies=out[0], entries=out[1], raw_candidates=out[2], manifolds=out[3], items=out[4])

    def set_profiling(self, on: bool):
        N.check(self.lib.r2d_ctx_set_profiling(self.handle, 1 if on else 0), "r2d_ctx_set_profiling")
        self.profiling = bool(on)

    def phase_times(self):
        ms = (ctypes.c_double * 4)()
        steps = ctypes.c_uint64()
        N.check(self.lib.r2d_ctx_phase_times(self.handle, ms, ctypes.byref(steps)), "r2d_ctx_phase_times")
        return list(ms), steps.value

    def step_times(self):
        """Device time (ms) of every timed step since the last reset."""
        n = ctypes.c_size_t()
        N.check(self.lib.r2d_ctx_step_times(self.handle, None, 0, ctypes.byref(n)), "r2d_ctx_step_times")
        out = (ctypes.c_float * max(n.value, 1))()
        N.check(self.lib.r2d_ctx_step_times(self.handle, out, n.value, ctypes.byref(n)), "r2d_ctx_step_times")
        return list(out)[:n.value]

    def measured_peaks(self):
        """The device's measured peaks in GB/s: HBM read stream, HBM copy
        (read + written bytes), LDS reads over all CUs."""
        out = (ctypes.c_double * 3)()
        N.check(self.lib.r2d_debug_peaks(self.handle, out), "r2d_debug_peaks")
        return {"hbm_read": out[0], "hbm_copy": out[1], "lds_read": out[2]}

    def reset_phase_times(self):
        N.check(self.lib.r2d_ctx_reset_phase_times(self.handle), "r2d_ctx_reset_phase_times")

    def last_counts(self):
        out = (ctypes.c_uint64 * 4)()
        N.check(self.lib.r2d_ctx_last_counts(self.handle, out), "r2d_ctx_last_counts")
        return dict(entries=out[0], raw_candidates=out[1], manifolds=out[2], items=out[3])


class PreparedBatch:
    """Worlds stepped together by r2d_process_batch, with the C array of
    their handles built once."""

    def __init__(self, ctx: Context, handles):
        self.ctx = ctx
        self.n = len(handles)
        self.arr = (ctypes.c_void_p * max(self.n, 1))(*handles)

    def process(self, dt: float, sub_steps: int, collision_iters: int, steps: int = 1):
        N.check(self.ctx.lib.r2d_process_batch(self.ctx.handle, self.arr, self.n, dt, sub_steps, collision_iters,
                                               steps), "r2d_process_batch")


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class BodyOptions:
    """EntityFactory.BodyOptions (lib.zig:35-46)."""
    pos: Vec2
    vel: Vec2 = (0.0, 0.0)
    angle: float = 0.0
    omega: float = 0.0
    mu: float = 0.5
    density: Optional[float] = None
    mass: Optional[float] = None

    def to_c(self):
        if (self.density is None) == (self.mass is None):
            raise ValueError("exactly one of density / mass")
        o = N.BodyOptions()
        o.pos_x, o.pos_y = self.pos
        o.vel_x, o.vel_y = self.vel
        o.angle, o.omega, o.mu = self.angle, self.omega, self.mu
        o.mass_kind = 0 if self.density is not None else 1
        o.mass_value = self.density if self.density is not None else self.mass
        return o


@dataclass
class DiscOptions:
    radius: float = 1.0


@dataclass
class RectangleOptions:
    width: float = 1.0
    height: float = 0.5


@dataclass
class Parameters:
    """Constraint.Parameters (Constraints/Constraint.zig:27-31)."""
    power_max: float = math.inf
    power_min: float = -math.inf
    beta: float = 10.0

    def to_c(self):
        p = N.ConstraintParams()
        p.power_max, p.power_min, p.beta = self.power_max, self.power_min, self.beta
        return p


class _Props:
    """RigidBody.Props view (RigidBody.zig:69-81)."""

    def __init__(self, body: "RigidBody"):
        object.__setattr__(self, "_b", body)

    _MAP = {"pos": ("pos_x", "pos_y"), "momentum": ("mom_x", "mom_y"), "force": ("force_x", "force_y")}
    _SCALARS = {"mass": "mass", "angle": "angle", "ang_momentum": "ang_momentum", "torque": "torque",
                "inertia": "inertia", "mu": "mu"}

    def __getattr__(self, name):
        st = self._b._state()
        if name in self._MAP:
            a, b = self._MAP[name]
            return (getattr(st, a), getattr(st, b))
        if name in self._SCALARS:
            return getattr(st, self._SCALARS[name])
        raise AttributeError(name)

    def __setattr__(self, name, value):
        st = self._b._state()
        if name in self._MAP:
            a, b = self._MAP[name]
            setattr(st, a, value[0])
            setattr(st, b, value[1])
        elif name in ("angle", "ang_momentum", "torque"):
            setattr(st, name, value)
        else:
            raise AttributeError(f"{name} is not writable between steps")
        self._b._write(st)


class RigidBody:
    """Live view of one body (RigidBody.zig:83-91)."""

    def __init__(self, solver: "Solver", id_: int):
        self._solver = solver
        self.id = id_

    def _state(self):
        st = N.BodyState()
        N.check(self._solver.lib.r2d_get_body(self._solver.handle, self.id, ctypes.byref(st)), "r2d_get_body")
        return st

    def _write(self, st):
        N.check(self._solver.lib.r2d_set_body(self._solver.handle, self.id, ctypes.byref(st)), "r2d_set_body")

    @property
    def props(self):
        return _Props(self)

    @property
    def static(self) -> bool:
        return bool(self._state().is_static)

    @static.setter
    def static(self, v: bool):
        N.check(self._solver.lib.r2d_set_static(self._solver.handle, self.id, 1 if v else 0), "r2d_set_static")

    @property
    def aabb(self):
        st = self._state()
        return (st.aabb_x, st.aabb_y, st.aabb_half_width, st.aabb_half_height)


class BodyHandle:
    """EntityFactory.BodyHandle (lib.zig:15-32)."""

    def __init__(self, solver: "Solver", id_: int):
        self.id = id_
        self.solver = solver

    def body(self) -> Optional[RigidBody]:
        n = ctypes.c_size_t()
        st = N.BodyState()
        rc = self.solver.lib.r2d_get_body(self.solver.handle, self.id, ctypes.byref(st))
        if rc == 2:
            return None
        N.check(rc, "r2d_get_body")
        del n
        return RigidBody(self.solver, self.id)

    def body_unwrap(self) -> RigidBody:
        b = self.body()
        if b is None:
            raise RuntimeError("Tried to unwrap a null *RigidBody")
        return b


class EntityFactory:
    """EntityFactory (lib.zig:13-130)."""

    def __init__(self, solver: "Solver"):
        self.solver = solver

    def makeDiscBody(self, bo: BodyOptions, go: DiscOptions = DiscOptions()) -> BodyHandle:
        i = ctypes.c_uint32()
        o = bo.to_c()
        N.check(self.solver.lib.r2d_make_disc(self.solver.handle, ctypes.byref(o), go.radius, ctypes.byref(i)),
                "r2d_make_disc")
        return BodyHandle(self.solver, i.value)

    def makeRectangleBody(self, bo: BodyOptions, go: RectangleOptions = RectangleOptions()) -> BodyHandle:
        i = ctypes.c_uint32()
        o = bo.to_c()
        N.check(self.solver.lib.r2d_make_rect(self.solver.handle, ctypes.byref(o), go.width, go.height,
                                              ctypes.byref(i)), "r2d_make_rect")
        return BodyHandle(self.solver, i.value)

    def makeDownwardsGravity(self, g: float):
        N.check(self.solver.lib.r2d_make_gravity(self.solver.handle, g), "r2d_make_gravity")

    def makeSpring(self, h: BodyHandle, k: float, c: float, anchor: Vec2):
        """Anchored spring force generator (extension; the reference has none)."""
        N.check(self.solver.lib.r2d_make_spring(self.solver.handle, h.id, k, c, anchor[0], anchor[1]),
                "r2d_make_spring")

    def _handle(self, rc, out, where):
        N.check(rc, where)
        return out.value

    def makeDistanceJoint(self, params: Parameters, h1: BodyHandle, h2: BodyHandle, target_distance: float) -> int:
        out = ctypes.c_size_t()
        p = params.to_c()
        return self._handle(self.solver.lib.r2d_make_distance_joint(self.solver.handle, ctypes.byref(p), h1.id,
                                                                    h2.id, target_distance, ctypes.byref(out)),
                            out, "r2d_make_distance_joint")

    def makeOffsetDistanceJoint(self, params: Parameters, h1: BodyHandle, h2: BodyHandle, r1: Vec2, r2: Vec2,
                                target_distance: float) -> int:
        out = ctypes.c_size_t()
        p = params.to_c()
        return self._handle(self.solver.lib.r2d_make_offset_distance_joint(
            self.solver.handle, ctypes.byref(p), h1.id, h2.id, r1[0], r1[1], r2[0], r2[1], target_distance,
            ctypes.byref(out)), out, "r2d_make_offset_distance_joint")

    def makeFixedPositionJoint(self, params: Parameters, h: BodyHandle, target_position: Vec2) -> int:
        out = ctypes.c_size_t()
        p = params.to_c()
        return self._handle(self.solver.lib.r2d_make_fixed_position_joint(
            self.solver.handle, ctypes.byref(p), h.id, target_position[0], target_position[1], ctypes.byref(out)),
            out, "r2d_make_fixed_position_joint")

    def makeMotorJoint(self, params: Parameters, h: BodyHandle, target_omega: float) -> int:
        out = ctypes.c_size_t()
        p = params.to_c()
        return self._handle(self.solver.lib.r2d_make_motor_joint(self.solver.handle, ctypes.byref(p), h.id,
                                                                 target_omega, ctypes.byref(out)),
                            out, "r2d_make_motor_joint")

    def excludeCollisionPair(self, h1: BodyHandle, h2: BodyHandle):
        N.check(self.solver.lib.r2d_exclude_collision_pair(self.solver.handle, h1.id, h2.id),
                "r2d_exclude_collision_pair")


class Solver:
    """Solver (lib.zig:132-315) with the step on the GPU."""

    def __init__(self, handle, ctx: Context):
        self.handle = handle
        self.ctx = ctx
        self.lib = ctx.lib

    @classmethod
    def init(cls, spatialhash_cell_width: float = 4.0, spatialhash_table_size_mult: int = 2,
             ctx: Optional[Context] = None) -> "Solver":
        ctx = ctx or default_context()
        h = ctypes.c_void_p()
        N.check(ctx.lib.r2d_world_create(ctx.handle, spatialhash_cell_width, spatialhash_table_size_mult,
                                         ctypes.byref(h)), "r2d_world_create")
        s = cls(h, ctx)
        ctx.adopt(s)
        return s

    def close(self):
        self.deinit()

    def deinit(self):
        if self.handle:
            N.check(self.lib.r2d_world_destroy(self.handle), "r2d_world_destroy")
            self.handle = None
            self.ctx.release(self)

    def clear(self):
        N.check(self.lib.r2d_world_clear(self.handle), "r2d_world_clear")

    def entityFactory(self) -> EntityFactory:
        return EntityFactory(self)

    def bodyHandle(self, id_: int) -> BodyHandle:
        return BodyHandle(self, id_)

    def removeRigidBody(self, id_: int):
        N.check(self.lib.r2d_remove_body(self.handle, id_), "r2d_remove_body")

    def process(self, dt: float, sub_steps: int, collision_iters: int):
        N.check(self.lib.r2d_process(self.handle, dt, sub_steps, collision_iters), "r2d_process")

    # wasm_root.zig-style accessors
    def num_bodies(self) -> int:
        n = ctypes.c_size_t()
        N.check(self.lib.r2d_num_bodies(self.handle, ctypes.byref(n)), "r2d_num_bodies")
        return n.value

    def body_id_at(self, i: int) -> int:
        out = ctypes.c_uint32()
        N.check(self.lib.r2d_body_id_at(self.handle, i, ctypes.byref(out)), "r2d_body_id_at")
        return out.value

    def states(self):
        n = self.num_bodies()
        arr = (N.BodyState * max(n, 1))()
        cnt = ctypes.c_size_t()
        N.check(self.lib.r2d_get_bodies(self.handle, arr, n, ctypes.byref(cnt)), "r2d_get_bodies")
        return arr[:n]

    def manifolds(self):
        cnt = ctypes.c_size_t()
        N.check(self.lib.r2d_get_manifolds(self.handle, None, 0, ctypes.byref(cnt)), "r2d_get_manifolds")
        arr = (N.ManifoldState * max(cnt.value, 1))()
        N.check(self.lib.r2d_get_manifolds(self.handle, arr, cnt.value, ctypes.byref(cnt)), "r2d_get_manifolds")
        return arr[:cnt.value]

    def set_hash_params(self, mode: int, cell_width: float = 4.0, table_mult: int = 2):
        N.check(self.lib.r2d_set_hash_params(self.handle, mode, cell_width, table_mult), "r2d_set_hash_params")

    def set_schedule(self, mode: int):
        N.check(self.lib.r2d_set_schedule(self.handle, mode), "r2d_set_schedule")

    def set_warm_start(self, factor: float):
        """Opt-in warm starting (r2d_set_warm_start): 0 off (the reference's
        results), (0, 1] the fraction of the previous step's impulses carried."""
        N.check(self.lib.r2d_set_warm_start(self.handle, factor), "r2d_set_warm_start")
