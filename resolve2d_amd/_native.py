"""ctypes binding of the C ABI declared in include/r2d.h.

The library is built in-tree by resolve2d_amd/build.py.  Nothing here falls
back to a CPU implementation: if the HIP library is missing, loading fails
with an ImportError that says how to build it.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libr2d_hip.so")

R2D_OK = 0
STATUS_NAMES = {
    0: "R2D_OK", 1: "R2D_E_INVALID_ARG", 2: "R2D_E_INVALID_ID", 3: "R2D_E_OOM",
    4: "R2D_E_NONFINITE", 5: "R2D_E_DEVICE", 6: "R2D_E_NO_DEVICE", 7: "R2D_E_RANGE",
    8: "R2D_E_CAPACITY", 9: "R2D_E_BUSY",
}

REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class R2DError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}")


class InvalidRigidBodyId(R2DError):
    """error.InvalidRigidBodyId / error.NoSuchIdExists of the reference."""


class BodyOptions(ctypes.Structure):
    _fields_ = [("pos_x", ctypes.c_float), ("pos_y", ctypes.c_float),
                ("vel_x", ctypes.c_float), ("vel_y", ctypes.c_float),
                ("angle", ctypes.c_float), ("omega", ctypes.c_float), ("mu", ctypes.c_float),
                ("mass_kind", ctypes.c_int32), ("mass_value", ctypes.c_float)]


class ConstraintParams(ctypes.Structure):
    _fields_ = [("power_max", ctypes.c_float), ("power_min", ctypes.c_float), ("beta", ctypes.c_float)]


class BodyState(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("type", ctypes.c_uint32), ("is_static", ctypes.c_uint32),
                ("num_normals", ctypes.c_uint32),
                ("pos_x", ctypes.c_float), ("pos_y", ctypes.c_float),
                ("mom_x", ctypes.c_float), ("mom_y", ctypes.c_float),
                ("force_x", ctypes.c_float), ("force_y", ctypes.c_float),
                ("mass", ctypes.c_float), ("angle", ctypes.c_float), ("ang_momentum", ctypes.c_float),
                ("torque", ctypes.c_float), ("inertia", ctypes.c_float), ("mu", ctypes.c_float),
                ("aabb_x", ctypes.c_float), ("aabb_y", ctypes.c_float),
                ("aabb_half_width", ctypes.c_float), ("aabb_half_height", ctypes.c_float),
                ("radius", ctypes.c_float), ("width", ctypes.c_float), ("height", ctypes.c_float)]


class ManifoldPoint(ctypes.Structure):
    _fields_ = [("valid", ctypes.c_uint32)] + [(n, ctypes.c_float) for n in (
        "acc_pn", "acc_pt", "ref_r_x", "ref_r_y", "inc_r_x", "inc_r_y", "pos_x", "pos_y",
        "depth", "original_depth", "mass_n", "mass_t")]


class ManifoldState(ctypes.Structure):
    _fields_ = [("ref_id", ctypes.c_uint32), ("inc_id", ctypes.c_uint32),
                ("normal_x", ctypes.c_float), ("normal_y", ctypes.c_float),
                ("tangent_x", ctypes.c_float), ("tangent_y", ctypes.c_float),
                ("friction", ctypes.c_float), ("points", ManifoldPoint * 2)]


class ServeOptions(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_worlds", ctypes.c_uint32), ("dt", ctypes.c_float),
                ("sub_steps", ctypes.c_uint32), ("collision_iters", ctypes.c_uint32),
                ("split_bodies", ctypes.c_uint32 * 2), ("steps_per_tick", ctypes.c_uint32)]


class ServeTickInfo(ctypes.Structure):
    _fields_ = [("admitted", ctypes.c_uint32), ("completed", ctypes.c_uint32), ("queued", ctypes.c_uint32),
                ("running", ctypes.c_uint32), ("launched_worlds", ctypes.c_uint32),
                ("launched_bodies", ctypes.c_uint64), ("resident_bytes", ctypes.c_size_t),
                ("step_gpu_ms", ctypes.c_double), ("step_wall_ms", ctypes.c_double)]


class RequestStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("steps_done", ctypes.c_uint32),
                ("steps_requested", ctypes.c_uint32), ("footprint_bytes", ctypes.c_size_t), ("lane", ctypes.c_uint32),
                ("wait_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("max_step_ms", ctypes.c_double)]


class ServeSummary(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "rejected", "queued", "running",
                                               "batch_steps", "scene_steps", "body_steps")] + \
               [(n, ctypes.c_double) for n in ("gpu_ms", "wall_ms", "p50_step_ms", "p99_step_ms", "p50_batch_ms",
                                               "p99_batch_ms")] + \
               [("resident_bytes", ctypes.c_size_t), ("peak_resident_bytes", ctypes.c_size_t),
                ("lanes", ctypes.c_uint32)]


P = ctypes.c_void_p
U32P = ctypes.POINTER(ctypes.c_uint32)
SZP = ctypes.POINTER(ctypes.c_size_t)
F = ctypes.c_float
SZ = ctypes.c_size_t
U32 = ctypes.c_uint32
I32 = ctypes.c_int

# name -> argtypes (every function returns int status)
SIGNATURES = {
    "r2d_ctx_create": [I32, ctypes.POINTER(P)],
    "r2d_ctx_destroy": [P],
    "r2d_ctx_sync": [P],
    "r2d_world_create": [P, F, SZ, ctypes.POINTER(P)],
    "r2d_world_destroy": [P],
    "r2d_world_clear": [P],
    "r2d_make_disc": [P, ctypes.POINTER(BodyOptions), F, U32P],
    "r2d_make_rect": [P, ctypes.POINTER(BodyOptions), F, F, U32P],
    "r2d_set_static": [P, U32, I32],
    "r2d_remove_body": [P, U32],
    "r2d_make_gravity": [P, F],
    "r2d_make_spring": [P, U32, F, F, F, F],
    "r2d_make_distance_joint": [P, ctypes.POINTER(ConstraintParams), U32, U32, F, SZP],
    "r2d_make_offset_distance_joint": [P, ctypes.POINTER(ConstraintParams), U32, U32, F, F, F, F, F, SZP],
    "r2d_make_fixed_position_joint": [P, ctypes.POINTER(ConstraintParams), U32, F, F, SZP],
    "r2d_make_motor_joint": [P, ctypes.POINTER(ConstraintParams), U32, F, SZP],
    "r2d_exclude_collision_pair": [P, U32, U32],
    "r2d_process": [P, F, SZ, SZ],
    "r2d_process_batch": [P, ctypes.POINTER(P), SZ, F, SZ, SZ, SZ],
    "r2d_process_static": [P, F, SZ, SZ],
    "r2d_num_bodies": [P, SZP],
    "r2d_body_id_at": [P, SZ, U32P],
    "r2d_get_body": [P, U32, ctypes.POINTER(BodyState)],
    "r2d_set_body": [P, U32, ctypes.POINTER(BodyState)],
    "r2d_get_bodies": [P, ctypes.POINTER(BodyState), SZ, SZP],
    "r2d_get_manifolds": [P, ctypes.POINTER(ManifoldState), SZ, SZP],
    "r2d_set_hash_params": [P, I32, F, SZ],
    "r2d_set_schedule": [P, I32],
    "r2d_set_warm_start": [P, F],
    "r2d_ctx_set_profiling": [P, I32],
    "r2d_ctx_phase_times": [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)],
    "r2d_ctx_reset_phase_times": [P],
    "r2d_ctx_step_times": [P, ctypes.POINTER(ctypes.c_float), SZ, SZP],
    "r2d_ctx_last_counts": [P, ctypes.POINTER(ctypes.c_uint64)],
    "r2d_debug_eval": [P, I32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), SZ],
    "r2d_debug_peaks": [P, ctypes.POINTER(ctypes.c_double)],
    "r2d_last_error": [ctypes.c_char_p, SZ],
    "r2d_ctx_set_solve_variant": [P, I32],
    "r2d_ctx_set_lds_limit": [P, U32],
    "r2d_ctx_set_fused": [P, I32],
    "r2d_ctx_set_wide_record": [P, I32],
    "r2d_ctx_fused_phase_times": [P, ctypes.POINTER(ctypes.c_double)],
    "r2d_ctx_fused_counts": [P, ctypes.POINTER(ctypes.c_uint64)],
    "r2d_ctx_fused_times": [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                            ctypes.POINTER(ctypes.c_uint64)],
    "r2d_make_bodies": [P, SZ, ctypes.POINTER(BodyOptions), ctypes.POINTER(ctypes.c_int32),
                        ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), U32P],
    "r2d_world_footprint": [P, SZP],
    "r2d_serve_create": [P, ctypes.POINTER(ServeOptions), ctypes.POINTER(P)],
    "r2d_serve_destroy": [P],
    "r2d_serve_submit": [P, P, U32, ctypes.POINTER(ctypes.c_uint64)],
    "r2d_serve_tick": [P, ctypes.POINTER(ServeTickInfo)],
    "r2d_serve_drain": [P],
    "r2d_serve_poll": [P, ctypes.c_uint64, ctypes.POINTER(RequestStatus)],
    "r2d_serve_completed": [P, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)],
    "r2d_serve_stats": [P, ctypes.POINTER(ServeSummary)],
}

_lib = None


def load():
    """Load the HIP library (once).  Raises ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python resolve2d_amd/build.py` "
                          "(or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = lib
    return lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    if _lib is not None:
        _lib.r2d_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(status, where):
    if status != R2D_OK:
        where = f"{where} [{last_error()}]"
        if status == 2:
            raise InvalidRigidBodyId(status, where)
        raise R2DError(status, where)
