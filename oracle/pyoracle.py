"""ORACLE (test infrastructure only): ctypes binding of oracle/build/libr2d_oracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker or as the timed CPU baseline.  `OracleWorld` offers the
same scene operations and array exports as resolve2d_amd.engine.GpuWorld.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libr2d_oracle.so")
BODY_COLS = 19
MANIFOLD_COLS = 33

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def build_planted(plant: int, out_dir: str) -> str:
    """Compile the oracle with the deliberate mistake number `plant`
    (R2D_PLANT in r2d_oracle_impl.h) into out_dir; tests use it to show
    that their checks catch that mistake."""
    path = os.path.join(out_dir, f"libr2d_oracle_plant{plant}.so")
    subprocess.run(["gcc", "-std=gnu11", "-O1", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    f"-DR2D_PLANT={plant}", "-o", path, os.path.join(HERE, "r2d_oracle.c"), "-lm"], check=True)
    return path


def load_path(path: str):
    """A separately built oracle library (see build_planted)."""
    return _bind(ctypes.CDLL(path))


def _bind(lib):
    D, U32, I, P, I64, U64 = ctypes.c_double, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64
    for pre in ("or32_", "or64_"):
        sig = {
            "world_new": (P, [D, U64]), "world_free_": (None, [P]), "set_hash": (None, [P, D, U64]),
            "clear": (None, [P]),
            "make_disc_": (U32, [P, D, D, D, D, D, D, D, I, D, D]),
            "make_rect_": (U32, [P, D, D, D, D, D, D, D, I, D, D, D]),
            "set_static": (I, [P, U32, I]), "remove_body_": (I, [P, U32]),
            "make_gravity": (None, [P, D]), "make_spring": (None, [P, U32, D, D, D, D]),
            "make_joint": (I64, [P, I, D, D, D, U32, U32, D, D, D, D, D, D, D]),
            "exclude_": (None, [P, U32, U32]), "process_": (I, [P, D, U64, U64]),
            "num_bodies": (I, [P]), "num_manifolds": (I, [P]), "current_id": (U32, [P]),
            "get_bodies": (None, [P, ctypes.POINTER(D)]),
            "set_body": (I, [P, U32, D, D, D, D, D, D, D, D, D]),
            "get_manifolds": (None, [P, ctypes.POINTER(D)]),
            "set_trace": (None, [P, I]), "num_trace": (I64, [P]), "set_schedule": (None, [P, I]),
            "get_trace": (None, [P, ctypes.POINTER(U32)]),
            "narrow_pair": (I, [P, U32, U32, ctypes.POINTER(D)]),
            "set_warm": (None, [P, D]), "num_warm": (I, [P]), "set_warm_factor": (None, [P, D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, pre + name)
            fn.restype = res
            fn.argtypes = args
    lib.or_sinf.restype = ctypes.c_float
    lib.or_sinf.argtypes = [ctypes.c_float]
    lib.or_cosf.restype = ctypes.c_float
    lib.or_cosf.argtypes = [ctypes.c_float]
    return lib


class OracleWorld:
    """Reference step restated in C; precision "f32" (bit-level reference) or "f64"."""

    def __init__(self, precision="f32", cell_width=4.0, table_mult=2, lib=None):
        self.lib = lib if lib is not None else load()
        self.p = "or32_" if precision == "f32" else "or64_"
        self.h = self._f("world_new")(cell_width, table_mult)

    def _f(self, name):
        return getattr(self.lib, self.p + name)

    def close(self):
        if self.h:
            self._f("world_free_")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def make_disc(self, px, py, vx, vy, angle, omega, mu, mass_kind, mass_val, radius):
        return self._f("make_disc_")(self.h, px, py, vx, vy, angle, omega, mu, mass_kind, mass_val, radius)

    def make_rect(self, px, py, vx, vy, angle, omega, mu, mass_kind, mass_val, width, height):
        return self._f("make_rect_")(self.h, px, py, vx, vy, angle, omega, mu, mass_kind, mass_val, width, height)

    def set_static(self, id_, flag):
        if self._f("set_static")(self.h, id_, 1 if flag else 0):
            raise KeyError(id_)

    def remove_body(self, id_):
        if self._f("remove_body_")(self.h, id_):
            raise KeyError(id_)

    def make_gravity(self, g):
        self._f("make_gravity")(self.h, g)

    def make_spring(self, id_, k, c, ax, ay):
        self._f("make_spring")(self.h, id_, k, c, ax, ay)

    def make_joint(self, type_, pmax, pmin, beta, id1, id2, r1x, r1y, r2x, r2y, target, tx, ty):
        return self._f("make_joint")(self.h, type_, pmax, pmin, beta, id1, id2, r1x, r1y, r2x, r2y, target, tx, ty)

    def exclude(self, id1, id2):
        self._f("exclude_")(self.h, id1, id2)

    def set_body(self, id_, px, py, mx, my, fx, fy, angle, L, torque):
        if self._f("set_body")(self.h, id_, px, py, mx, my, fx, fy, angle, L, torque):
            raise KeyError(id_)

    def set_hash(self, cell, mult):
        self._f("set_hash")(self.h, cell, mult)

    def clear(self):
        self._f("clear")(self.h)

    def set_schedule(self, mode):
        self._f("set_schedule")(self.h, mode)

    def set_warm_start(self, factor):
        """Warm starting (an extension, not the reference's): 0 off."""
        self._f("set_warm")(self.h, factor)

    def set_warm_factor(self, factor):
        """A new factor, the kept manifolds left in place (tests only)."""
        self._f("set_warm_factor")(self.h, factor)

    def num_warm(self):
        """Manifolds the last step warm started."""
        return self._f("num_warm")(self.h)

    def process(self, dt, sub_steps, iters):
        return self._f("process_")(self.h, dt, sub_steps, iters)

    def bodies(self) -> np.ndarray:
        n = self._f("num_bodies")(self.h)
        out = np.zeros((n, BODY_COLS))
        if n:
            self._f("get_bodies")(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def manifolds(self) -> np.ndarray:
        n = self._f("num_manifolds")(self.h)
        out = np.zeros((n, MANIFOLD_COLS))
        if n:
            self._f("get_manifolds")(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def trace(self, on=True):
        self._f("set_trace")(self.h, 1 if on else 0)

    def traced_pairs(self) -> np.ndarray:
        n = self._f("num_trace")(self.h)
        out = np.zeros(2 * n, dtype=np.uint32)
        if n:
            self._f("get_trace")(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        return out.reshape(n, 2)

    def narrow_pair(self, id1, id2):
        out = np.zeros(3 + MANIFOLD_COLS)
        rc = self._f("narrow_pair")(self.h, id1, id2, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc < 0:
            raise KeyError((id1, id2))
        return out


def sinf(x):
    return load().or_sinf(x)


def cosf(x):
    return load().or_cosf(x)
