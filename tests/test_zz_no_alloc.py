"""r2d_process_static -- the step without allocation (its scratch is a static
device array) -- against r2d_process, the entry point it is a variant of,
called through the C ABI: bit-identical bodies and manifolds on calls in a
row, the second with a smaller world; the limits and preconditions it
states are refused with their errors."""
import ctypes

import numpy as np
import pytest

from helpers import bits_equal, first_diff

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _static(w):
    return w.lib.r2d_process_static(w.handle, DT, 4, 4)


def _pair(ctx, ops):
    from resolve2d_amd.engine import GpuWorld, apply_scene
    a, b = GpuWorld(ctx), GpuWorld(ctx)
    apply_scene(ops, a)
    apply_scene(ops, b)
    # resident: one ordinary step each
    assert a.process(DT, 4, 4) == 0 and b.process(DT, 4, 4) == 0
    return a, b


def _same(a, b, where):
    ba, bb = a.bodies(), b.bodies()
    assert bits_equal(ba, bb), (where, first_diff(ba, bb))
    ma, mb = a.manifolds(), b.manifolds()
    assert ma.shape == mb.shape and bits_equal(ma, mb), (where, ma.shape, mb.shape)


def test_static_step_equals_process(gpu_ctx):
    """Calls in a row on a larger then a smaller world (the shared scratch
    keeps the previous call's contents), then again, each against the same
    world stepped by r2d_process."""
    from resolve2d_amd import scenes
    big = _pair(gpu_ctx, scenes.batched_world(31, n_min=400, n_max=400))
    small = _pair(gpu_ctx, scenes.batched_world(32, n_min=60, n_max=60))
    for rnd in range(3):
        for (s, r), name in ((big, "400 bodies"), (small, "60 bodies")):
            assert _static(s) == 0, (rnd, name)
            assert r.process(DT, 4, 4) == 0
            _same(s, r, (rnd, name))
    for w in big + small:
        w.close()


def test_static_step_scenes_and_edits(gpu_ctx):
    """The pyramid and the car (all four joint kinds) over 30 calls each,
    a state edit between calls, and the oracle-checked reference path
    giving the same bits."""
    from resolve2d_amd import scenes
    for ops in (scenes.pyramid(rows=10), scenes.car_demo()):
        s, r = _pair(gpu_ctx, ops)
        for k in range(30):
            if k == 10:
                for w in (s, r):
                    w.set_body(2, 0.5, 3.0, 0.25, -0.5, 0.0, 0.0, 0.1, 0.3, 0.0)
            assert _static(s) == 0
            assert r.process(DT, 4, 4) == 0
        _same(s, r, "after 30 calls")
        s.close()
        r.close()


def test_static_step_refusals(gpu_ctx):
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld, apply_scene
    # not resident: bodies added since its last step
    w = GpuWorld(gpu_ctx)
    apply_scene(scenes.pyramid(rows=4), w)
    assert _static(w) == 1                      # R2D_E_INVALID_ARG
    assert w.process(DT, 4, 4) == 0
    assert _static(w) == 0
    w.close()
    # more bodies than the static scratch holds
    big = GpuWorld(gpu_ctx)
    apply_scene(scenes.mixed(1100, seed=2, width=60.0, height=60.0, ground_w=70.0), big)
    assert big.process(DT, 4, 4) == 0
    assert _static(big) == 8                    # R2D_E_CAPACITY
    big.close()
    # the coloured schedule is not offered
    c = GpuWorld(gpu_ctx)
    apply_scene(scenes.pyramid(rows=4), c)
    c.set_schedule(1)
    assert c.process(DT, 4, 4) == 0
    assert _static(c) == 1
    c.close()


def test_manifolds_overwritten_by_next_call(gpu_ctx):
    from resolve2d_amd import scenes
    a, ra = _pair(gpu_ctx, scenes.pyramid(rows=6))
    b, rb = _pair(gpu_ctx, scenes.pyramid(rows=3))
    assert _static(a) == 0 and ra.process(DT, 4, 4) == 0
    _same(a, ra, "a")
    assert _static(b) == 0
    n = ctypes.c_size_t()
    assert a.lib.r2d_get_manifolds(a.handle, None, 0, ctypes.byref(n)) == 7   # R2D_E_RANGE
    assert bits_equal(a.bodies(), ra.bodies())                               # bodies stay readable
    for w in (a, ra, b, rb):
        w.close()
