"""Scene application over the C ABI with plain numbers.

A scene is a list of operations (see resolve2d_amd.scenes); `GpuWorld`
applies them to one device world and exports state as arrays laid out like
the reference's body / manifold records:

  body row (19):     id, type, static, pos.x, pos.y, momentum.x, momentum.y,
                     force.x, force.y, mass, angle, ang_momentum, torque,
                     inertia, mu, aabb.x, aabb.y, aabb.half_width, aabb.half_height
  manifold row (33): ref id, inc id, normal.x/y, tangent.x/y, friction, then
                     two points of (valid, acc_pn, acc_pt, ref_r.x/y, inc_r.x/y,
                     pos.x/y, depth, original_depth, mass_n, mass_t)
"""
import ctypes

import numpy as np

from . import _native as N
from .solver import Context, default_context

BODY_COLS = 19
MANIFOLD_COLS = 33


class GpuWorld:
    def __init__(self, ctx: Context = None, cell_width: float = 4.0, table_mult: int = 2):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        h = ctypes.c_void_p()
        N.check(self.lib.r2d_world_create(self.ctx.handle, cell_width, table_mult, ctypes.byref(h)),
                "r2d_world_create")
        self.handle = h
        self.ctx.adopt(self)

    def close(self):
        if self.handle:
            N.check(self.lib.r2d_world_destroy(self.handle), "r2d_world_destroy")
            self.handle = None
            self.ctx.release(self)

    # --- scene operations (same names/arguments as the oracle's engine)
    def _opts(self, px, py, vx, vy, angle, omega, mu, mass_kind, mass_val):
        o = N.BodyOptions()
        o.pos_x, o.pos_y, o.vel_x, o.vel_y = px, py, vx, vy
        o.angle, o.omega, o.mu = angle, omega, mu
        o.mass_kind, o.mass_value = mass_kind, mass_val
        return o

    def make_disc(self, px, py, vx, vy, angle, omega, mu, mass_kind, mass_val, radius):
        o = self._opts(px, py, vx, vy, angle, omega, mu, mass_kind, mass_val)
        i = ctypes.c_uint32()
        N.check(self.lib.r2d_make_disc(self.handle, ctypes.byref(o), radius, ctypes.byref(i)), "r2d_make_disc")
        return i.value

    def make_rect(self, px, py, vx, vy, angle, omega, mu, mass_kind, mass_val, width, height):
        o = self._opts(px, py, vx, vy, angle, omega, mu, mass_kind, mass_val)
        i = ctypes.c_uint32()
        N.check(self.lib.r2d_make_rect(self.handle, ctypes.byref(o), width, height, ctypes.byref(i)),
                "r2d_make_rect")
        return i.value

    def set_static(self, id_, flag):
        N.check(self.lib.r2d_set_static(self.handle, id_, 1 if flag else 0), "r2d_set_static")

    def remove_body(self, id_):
        N.check(self.lib.r2d_remove_body(self.handle, id_), "r2d_remove_body")

    def make_gravity(self, g):
        N.check(self.lib.r2d_make_gravity(self.handle, g), "r2d_make_gravity")

    def make_spring(self, id_, k, c, ax, ay):
        N.check(self.lib.r2d_make_spring(self.handle, id_, k, c, ax, ay), "r2d_make_spring")

    def make_joint(self, type_, pmax, pmin, beta, id1, id2, r1x, r1y, r2x, r2y, target, tx, ty):
        p = N.ConstraintParams()
        p.power_max, p.power_min, p.beta = pmax, pmin, beta
        out = ctypes.c_size_t()
        if type_ == 0:
            rc = self.lib.r2d_make_distance_joint(self.handle, ctypes.byref(p), id1, id2, target, ctypes.byref(out))
        elif type_ == 1:
            rc = self.lib.r2d_make_offset_distance_joint(self.handle, ctypes.byref(p), id1, id2, r1x, r1y, r2x, r2y,
                                                         target, ctypes.byref(out))
        elif type_ == 2:
            rc = self.lib.r2d_make_fixed_position_joint(self.handle, ctypes.byref(p), id1, tx, ty, ctypes.byref(out))
        else:
            rc = self.lib.r2d_make_motor_joint(self.handle, ctypes.byref(p), id1, target, ctypes.byref(out))
        N.check(rc, "r2d_make_joint")
        return out.value

    def exclude(self, id1, id2):
        N.check(self.lib.r2d_exclude_collision_pair(self.handle, id1, id2), "r2d_exclude_collision_pair")

    def set_body(self, id_, px, py, mx, my, fx, fy, angle, L, torque):
        st = N.BodyState()
        N.check(self.lib.r2d_get_body(self.handle, id_, ctypes.byref(st)), "r2d_get_body")
        st.pos_x, st.pos_y, st.mom_x, st.mom_y = px, py, mx, my
        st.force_x, st.force_y, st.angle, st.ang_momentum, st.torque = fx, fy, angle, L, torque
        N.check(self.lib.r2d_set_body(self.handle, id_, ctypes.byref(st)), "r2d_set_body")

    def set_hash(self, cell, mult):
        N.check(self.lib.r2d_set_hash_params(self.handle, 1, cell, mult), "r2d_set_hash_params")

    def set_schedule(self, mode):
        N.check(self.lib.r2d_set_schedule(self.handle, mode), "r2d_set_schedule")

    def set_warm_start(self, factor):
        N.check(self.lib.r2d_set_warm_start(self.handle, factor), "r2d_set_warm_start")

    def process(self, dt, sub_steps, iters):
        return self.lib.r2d_process(self.handle, dt, sub_steps, iters)

    # --- state export
    def make_bodies(self, kind, px, py, vx, vy, angle, omega, mu, mass_kind, mass_value, dim_a, dim_b):
        """n bodies in one call (r2d_make_bodies); every argument an array of n."""
        n = len(kind)
        opts = (N.BodyOptions * max(n, 1))()
        rec = np.zeros(n, dtype=[("pos_x", "<f4"), ("pos_y", "<f4"), ("vel_x", "<f4"), ("vel_y", "<f4"),
                                 ("angle", "<f4"), ("omega", "<f4"), ("mu", "<f4"), ("mass_kind", "<i4"),
                                 ("mass_value", "<f4")])
        for name, v in (("pos_x", px), ("pos_y", py), ("vel_x", vx), ("vel_y", vy), ("angle", angle),
                        ("omega", omega), ("mu", mu), ("mass_kind", mass_kind), ("mass_value", mass_value)):
            rec[name] = v
        ctypes.memmove(opts, rec.ctypes.data, rec.nbytes)
        k = np.ascontiguousarray(kind, dtype=np.int32)
        a = np.ascontiguousarray(dim_a, dtype=np.float32)
        b = np.ascontiguousarray(dim_b, dtype=np.float32)
        ids = np.zeros(max(n, 1), dtype=np.uint32)
        fp = ctypes.POINTER(ctypes.c_float)
        N.check(self.lib.r2d_make_bodies(self.handle, n, opts, k.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                         a.ctypes.data_as(fp), b.ctypes.data_as(fp),
                                         ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))), "r2d_make_bodies")
        return ids[:n]

    def footprint(self) -> int:
        b = ctypes.c_size_t()
        N.check(self.lib.r2d_world_footprint(self.handle, ctypes.byref(b)), "r2d_world_footprint")
        return b.value

    def bodies(self) -> np.ndarray:
        n = ctypes.c_size_t()
        N.check(self.lib.r2d_num_bodies(self.handle, ctypes.byref(n)), "r2d_num_bodies")
        arr = (N.BodyState * max(n.value, 1))()
        N.check(self.lib.r2d_get_bodies(self.handle, arr, n.value, ctypes.byref(n)), "r2d_get_bodies")
        words = ctypes.sizeof(N.BodyState) // 4
        raw = np.frombuffer(arr, dtype=np.uint32, count=n.value * words).reshape(n.value, words)
        out = np.zeros((n.value, BODY_COLS))
        out[:, 0:3] = raw[:, 0:3]
        out[:, 3:] = raw[:, 4:20].view(np.float32)
        return out

    def manifolds(self) -> np.ndarray:
        cnt = ctypes.c_size_t()
        N.check(self.lib.r2d_get_manifolds(self.handle, None, 0, ctypes.byref(cnt)), "r2d_get_manifolds")
        arr = (N.ManifoldState * max(cnt.value, 1))()
        N.check(self.lib.r2d_get_manifolds(self.handle, arr, cnt.value, ctypes.byref(cnt)), "r2d_get_manifolds")
        m = cnt.value
        words = ctypes.sizeof(N.ManifoldState) // 4            # 2 ids, 5 floats, 2 x (valid + 12 floats)
        raw = np.frombuffer(arr, dtype=np.uint32, count=m * words).reshape(m, words)
        out = np.zeros((m, MANIFOLD_COLS))
        out[:, 0:2] = raw[:, 0:2]
        out[:, 2:7] = raw[:, 2:7].view(np.float32)
        for p in range(2):
            base = 7 + 13 * p
            valid = raw[:, base] != 0
            out[:, base] = valid
            vals = raw[:, base + 1:base + 13].view(np.float32)
            out[:, base + 1:base + 13] = np.where(valid[:, None], vals, 0.0)
        return out


def apply_scene(scene, eng):
    """Apply scene operations to an engine; returns the ids of created bodies."""
    ids = []
    for op in scene:
        kind = op[0]
        if kind == "disc":
            ids.append(eng.make_disc(*op[1:]))
        elif kind == "rect":
            ids.append(eng.make_rect(*op[1:]))
        elif kind == "static":
            eng.set_static(op[1], op[2])
        elif kind == "remove":
            eng.remove_body(op[1])
        elif kind == "gravity":
            eng.make_gravity(op[1])
        elif kind == "spring":
            eng.make_spring(*op[1:])
        elif kind == "joint":
            eng.make_joint(*op[1:])
        elif kind == "exclude":
            eng.exclude(op[1], op[2])
        elif kind == "hash":
            eng.set_hash(op[1], op[2])
        else:
            raise ValueError(kind)
    return ids
