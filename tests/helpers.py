"""Shared helpers: bit-level comparison of engine states."""
import numpy as np


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    """Exact equality of f32 values carried as f64, incl. the sign of zero;
    NaNs compare equal to NaNs."""
    if a.shape != b.shape:
        return False
    a32 = a.astype(np.float32)
    b32 = b.astype(np.float32)
    both_nan = np.isnan(a32) & np.isnan(b32)
    same = (a32.view(np.uint32) == b32.view(np.uint32)) | both_nan
    return bool(same.all())


def first_diff(a: np.ndarray, b: np.ndarray):
    a32 = a.astype(np.float32)
    b32 = b.astype(np.float32)
    ne = (a32.view(np.uint32) != b32.view(np.uint32)) & ~(np.isnan(a32) & np.isnan(b32))
    idx = np.argwhere(ne)
    if len(idx) == 0:
        return None
    r, c = idx[0]
    return (int(r), int(c), float(a[r, c]), float(b[r, c]), int(len(idx)))


def run_pair(scene, eng_a, eng_b, steps, dt=1.0 / 60.0, sub=4, iters=4, check_every=1):
    """Step two engines in lockstep; return the first step whose body state differs."""
    from resolve2d_amd.engine import apply_scene
    apply_scene(scene, eng_a)
    apply_scene(scene, eng_b)
    if not bits_equal(eng_a.bodies(), eng_b.bodies()):
        return 0, first_diff(eng_a.bodies(), eng_b.bodies())
    for s in range(1, steps + 1):
        ra = eng_a.process(dt, sub, iters)
        rb = eng_b.process(dt, sub, iters)
        assert ra == 0 and rb == 0, (s, ra, rb)
        if s % check_every == 0 or s == steps:
            a, b = eng_a.bodies(), eng_b.bodies()
            if not bits_equal(a, b):
                return s, first_diff(a, b)
    return None, None
