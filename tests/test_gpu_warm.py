"""Warm starting on the GPU (opt-in, r2d_set_warm_start) against the oracle's
warm starting, bit for bit: the fused kernel and the phase kernels, worlds
beyond LDS (global-memory state), warm and cold worlds in one batch, kept
manifolds across calls, removals and a new factor between steps.  The
kernels of cold batches carry no warm-start code; batches with a warm world
take the warm instantiation, which must leave its cold worlds' bits as the
reference's."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff
from resolve2d_amd import scenes
from resolve2d_amd.engine import GpuWorld, apply_scene

pytestmark = pytest.mark.gpu

DT = scenes.DT


def _pair(ctx, oracle, ops, warm):
    g, o = GpuWorld(ctx), oracle.OracleWorld("f32")
    for e in (g, o):
        apply_scene(ops, e)
        if warm:
            e.set_warm_start(warm)
    return g, o


def _same(g, o, what):
    a, b = g.bodies(), o.bodies()
    assert bits_equal(a, b), (what, first_diff(a, b))
    mg, mo = g.manifolds(), o.manifolds()
    assert mg.shape == mo.shape and bits_equal(mg, mo), (what, "manifolds", first_diff(mg, mo))


def _batch(ctx, oracle, specs, warms, calls):
    """specs stepped together over `calls` (steps per call), the oracle world
    by world; every world compared after every call."""
    pairs = [_pair(ctx, oracle, ops, w) for ops, w in zip(specs, warms)]
    for k, steps in enumerate(calls):
        print(f"warm batch: {len(specs)} worlds, call {k} of {steps} steps", flush=True)
        ctx.process_batch_handles([g.handle for g, _ in pairs], DT, 4, 4, steps)
        for i, (g, o) in enumerate(pairs):
            for _ in range(steps):
                assert o.process(DT, 4, 4) == 0
            _same(g, o, (i, k))
    for g, o in pairs:
        g.close()
        o.close()


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "phases"])
def test_warm_batch_matches_oracle(oracle, fused):
    from resolve2d_amd.solver import Context
    ctx = Context(0)
    ctx.set_fused(fused)
    specs = [scenes.batched_world(500 + i) for i in range(8)] + [scenes.pyramid(rows=12)]
    warms = [0.8, 0.0, 1.0, 0.8, 0.0, 0.5, 0.8, 0.0, 0.8]
    _batch(ctx, oracle, specs, warms, calls=(1, 14, 1, 24))
    ctx.close()


def test_warm_beyond_lds_matches_oracle(oracle):
    """Worlds whose body state exceeds the LDS limit (set to 64 bodies) run the
    grid-wide body / item kernels and the global-memory or compact chains."""
    from resolve2d_amd.solver import Context
    ctx = Context(0)
    ctx.set_fused(False)
    ctx.set_lds_limit(64)
    specs = [scenes.batched_world(600 + i) for i in range(3)] + [scenes.pyramid(rows=10)]
    _batch(ctx, oracle, specs, [0.8, 1.0, 0.0, 0.8], calls=(12, 8))
    ctx.close()


def test_warm_removal_factor_and_setters(gpu_ctx, oracle):
    g, o = _pair(gpu_ctx, oracle, scenes.mixed(300, seed=21, width=20.0, height=30.0, ground_w=30.0), 0.8)
    print("warm world: 300 bodies, removal / factor / setters between steps", flush=True)
    for phase in range(4):
        for _ in range(15):
            assert g.process(DT, 4, 4) == 0 and o.process(DT, 4, 4) == 0
        _same(g, o, phase)
        for e in (g, o):
            if phase == 0:
                e.remove_body(17)
            elif phase == 1:
                e.set_warm_start(1.0)
            elif phase == 2:
                e.set_body(40, 1.0, 8.0, 0.0, -3.0, 0.0, 0.0, 0.2, 0.0, 0.0)
    assert o.num_warm() > 0
    g.close()
    o.close()


def test_warm_is_refused_with_the_coloured_schedule(gpu_ctx):
    from resolve2d_amd._native import R2DError
    w = GpuWorld(gpu_ctx)
    apply_scene(scenes.pyramid(rows=3), w)
    w.set_schedule(1)
    with pytest.raises(R2DError):
        w.set_warm_start(0.8)
    w.set_schedule(0)
    w.set_warm_start(0.8)
    with pytest.raises(R2DError):
        w.set_schedule(1)
    for bad in (-0.1, 1.5, float("nan")):
        with pytest.raises(R2DError):
            w.set_warm_start(bad)
    w.close()


def test_warm_tables_survive_growth(gpu_ctx, oracle):
    """Bodies added between steps grow the device body block past its
    capacity: the warm tables move to larger ones and keep what they held,
    so the next step still carries impulses, as in the oracle."""
    from resolve2d_amd.scenes import _Builder
    base = _Builder()
    base.gravity(9.81)
    base.static(base.rect(0.0, -0.5, 40.0, 1.0))
    for i in range(6):
        for j in range(3):
            base.rect(-6.0 + 2.0 * i, 0.5 + 1.0 * j, 1.0, 1.0)
    more = _Builder()
    for i in range(40):
        more.disc(-15.0 + 0.8 * i, 8.0 + (i % 3), 0.35)
    g, o = _pair(gpu_ctx, oracle, base.ops, 0.8)
    print("warm world: 19 bodies, 40 added after 8 steps", flush=True)
    for _ in range(8):
        assert g.process(DT, 4, 4) == 0 and o.process(DT, 4, 4) == 0
    _same(g, o, "before growth")
    for e in (g, o):
        apply_scene(more.ops, e)
    assert g.process(DT, 4, 4) == 0 and o.process(DT, 4, 4) == 0
    assert o.num_warm() > 0
    for _ in range(20):
        assert g.process(DT, 4, 4) == 0 and o.process(DT, 4, 4) == 0
    _same(g, o, "after growth")
    g.close()
    o.close()


def test_warm_edge_steps_and_worlds(gpu_ctx, oracle):
    """Warm-started worlds through the step shapes the reference allows: no
    sub-steps (nothing carried is applied), no iterations (only the carried
    impulses act), one sub-step; and worlds with no bodies or only static
    ones.  Every step is compared with the oracle's warm starting."""
    from resolve2d_amd.scenes import _Builder
    shapes = [(4, 4), (4, 4), (0, 4), (4, 0), (1, 1), (4, 4), (2, 3), (4, 4)]
    g, o = _pair(gpu_ctx, oracle, scenes.mixed(120, seed=33, width=10.0, height=20.0, ground_w=20.0), 0.9)
    print("warm world: 120 bodies through sub-step / iteration edge cases", flush=True)
    for _ in range(10):
        assert g.process(DT, 4, 4) == 0 and o.process(DT, 4, 4) == 0
    for sub, it in shapes:
        assert g.process(DT, sub, it) == 0 and o.process(DT, sub, it) == 0
        _same(g, o, (sub, it))
    g.close()
    o.close()
    static_only = _Builder()
    static_only.gravity(9.81)
    static_only.static(static_only.rect(0.0, 0.0, 5.0, 1.0))
    static_only.static(static_only.rect(0.0, 0.9, 1.0, 1.0))
    for ops in ([], static_only.ops):
        g, o = _pair(gpu_ctx, oracle, ops, 0.8)
        for _ in range(3):
            assert g.process(DT, 4, 4) == 0 and o.process(DT, 4, 4) == 0
        _same(g, o, ("static or empty", len(ops)))
        g.close()
        o.close()
