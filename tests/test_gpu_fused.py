"""Whole steps per workgroup (the fused kernel).

A batch of worlds of at most 1,024 bodies in the reference order is stepped by
one kernel per call: every step of every world, one workgroup per world, the
broadphase, narrowphase, schedule and solve of a step separated by workgroup
barriers.  It must give the bits of the phase kernels (one launch sequence per
step) and of the oracle.  Also covered: the overflow protocol -- a world whose
scratch slice is too small stops before its state changes, the host grows the
slice and the same call resumes the world where it stopped, while the
manifolds of worlds that had already finished stay readable."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _worlds(ctx, specs):
    from resolve2d_amd.engine import GpuWorld, apply_scene
    out = []
    for ops in specs:
        w = GpuWorld(ctx)
        apply_scene(ops, w)
        out.append(w)
    return out


def _oracle_run(oracle, ops, steps):
    from resolve2d_amd.engine import apply_scene
    o = oracle.OracleWorld("f32")
    apply_scene(ops, o)
    for _ in range(steps):
        assert o.process(DT, 4, 4) == 0
    return o


def _on_block(block_w, block_h, seed):
    """40 small bodies falling onto a static block_w x block_h block beside a
    100 m ground: the block spans many hash cells, so the world needs more
    cell entries than the fused kernel's first estimate (4 per body + 64)."""
    from resolve2d_amd import scenes
    rng = np.random.default_rng(seed)
    b = scenes._Builder()
    b.gravity(9.81)
    b.static(b.rect(0.0, -0.5, 100.0, 1.0, mu=0.4))
    b.static(b.rect(0.0, block_h / 2, block_w, block_h, mu=0.4))
    half = block_w / 2 - 2.0
    for i in range(40):
        x, y = rng.uniform(-half, half), rng.uniform(block_h + 0.6, block_h + 4.0)
        if i % 2:
            b.disc(x, y, rng.uniform(0.2, 0.6), mu=0.4)
        else:
            b.rect(x, y, rng.uniform(0.3, 1.0), rng.uniform(0.3, 1.0), mu=0.4, angle=rng.uniform(0.0, 6.28))
    return b.ops


def _check_oracle(oracle, ws, specs, steps):
    for i, ops in enumerate(specs):
        o = _oracle_run(oracle, ops, steps)
        assert bits_equal(ws[i].bodies(), o.bodies()), (i, first_diff(ws[i].bodies(), o.bodies()))
        mg, mo = ws[i].manifolds(), o.manifolds()
        assert mg.shape == mo.shape and bits_equal(mg, mo), (i, first_diff(mg, mo))
        o.close()


def _launches(ctx, handles, steps):
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    ctx.process_batch_handles(handles, DT, 4, 4, steps)
    _, launches, world_steps = ctx.fused_times()
    ctx.set_profiling(False)
    return launches, world_steps


def test_fused_batch_equals_phase_kernels_and_oracle(oracle):
    from resolve2d_amd import scenes
    from resolve2d_amd.solver import Context
    steps, nw = 30, 24
    specs = [scenes.batched_world(100 + i) for i in range(nw)]
    fused, phase = Context(0), Context(0)
    fused.set_fused(True)
    phase.set_fused(False)
    a, b = _worlds(fused, specs), _worlds(phase, specs)
    print(f"fused batch: {nw} worlds x {steps} steps in one call", flush=True)
    fused.set_profiling(True)
    fused.reset_phase_times()
    fused.process_batch_handles([w.handle for w in a], DT, 4, 4, steps)
    ms, launches, world_steps = fused.fused_times()
    fused.set_profiling(False)
    assert launches == 1 and world_steps == nw * steps and ms > 0.0
    assert len(fused.step_times()) == nw * steps
    for _ in range(steps):
        phase.process_batch_handles([w.handle for w in b], DT, 4, 4, 1)
    for i in range(nw):
        ba, bb = a[i].bodies(), b[i].bodies()
        assert bits_equal(ba, bb), (i, first_diff(ba, bb))
        ma, mb = a[i].manifolds(), b[i].manifolds()
        assert ma.shape == mb.shape and bits_equal(ma, mb), (i, first_diff(ma, mb))
    for i in (0, 13):
        o = _oracle_run(oracle, specs[i], steps)
        assert bits_equal(a[i].bodies(), o.bodies()), (i, first_diff(a[i].bodies(), o.bodies()))
        assert bits_equal(a[i].manifolds(), o.manifolds())
        o.close()
    for w in a + b:
        w.close()
    fused.close()
    phase.close()


def test_fused_slice_grows_in_kernel(oracle):
    """The world on an 80 x 40 m block (about 360 cell entries for 232
    reserved) moves to a larger slice from the free end of the arrays inside
    the launch: one launch, every world equal to the oracle."""
    from resolve2d_amd import scenes
    from resolve2d_amd.solver import Context
    steps = 12
    specs = [scenes.batched_world(200 + i) for i in range(5)]
    specs.insert(2, _on_block(80.0, 40.0, 7))
    ctx = Context(0)
    ctx.set_fused(True)
    ws = _worlds(ctx, specs)
    print(f"fused slice growth: {len(specs)} worlds x {steps} steps", flush=True)
    launches, world_steps = _launches(ctx, [w.handle for w in ws], steps)
    assert launches == 1 and world_steps == len(specs) * steps
    _check_oracle(oracle, ws, specs, steps)
    for w in ws:
        w.close()
    ctx.close()


def test_fused_overflow_resumes(oracle):
    """The world on a 160 x 80 m block needs about 1,000 cell entries, more
    than the arrays' whole free end: it stops before its state changes, the
    host grows the arrays (keeping the other world's manifolds in place) and
    relaunches it for its remaining steps.  The next call fits at once."""
    from resolve2d_amd.solver import Context
    steps = 12
    specs = [_on_block(80.0, 40.0, 8), _on_block(160.0, 80.0, 9)]
    ctx = Context(0)
    ctx.set_fused(True)
    ws = _worlds(ctx, specs)
    print(f"fused overflow: {len(specs)} worlds x {steps} steps", flush=True)
    launches, world_steps = _launches(ctx, [w.handle for w in ws], steps)
    assert launches >= 2, "the world on the large block should have stopped for larger arrays"
    assert world_steps == len(specs) * steps
    _check_oracle(oracle, ws, specs, steps)
    assert _launches(ctx, [w.handle for w in ws], 2)[0] == 1
    for w in ws:
        w.close()
    ctx.close()


def test_fused_single_steps_equal_one_call():
    """Twenty one-step calls and one twenty-step call give the same bits."""
    from resolve2d_amd import scenes
    from resolve2d_amd.solver import Context
    specs = [scenes.batched_world(300 + i) for i in range(6)] + [scenes.pyramid(rows=10), scenes.car_demo()]
    ctx = Context(0)
    ctx.set_fused(True)
    a, b = _worlds(ctx, specs), _worlds(ctx, specs)
    ctx.process_batch_handles([w.handle for w in a], DT, 4, 4, 20)
    for _ in range(20):
        ctx.process_batch_handles([w.handle for w in b], DT, 4, 4, 1)
    for i in range(len(specs)):
        assert bits_equal(a[i].bodies(), b[i].bodies()), (i, first_diff(a[i].bodies(), b[i].bodies()))
    for w in a + b:
        w.close()
    ctx.close()


def test_size_boundary_of_the_fused_path(gpu_ctx, oracle):
    """Worlds of exactly 1,024 bodies (the fused kernel's limit, ground
    included) take the fused kernel; a batch holding one world of 1,025
    bodies takes the phase kernels for all its worlds.  Both give the
    oracle's bits over 20 steps."""
    from resolve2d_amd import scenes
    at = [scenes.mixed(1023, seed=s, width=30.0, height=40.0, ground_w=40.0) for s in (11, 12)]
    over = scenes.mixed(1024, seed=13, width=30.0, height=40.0, ground_w=40.0)
    small = scenes.mixed(60, seed=14, width=10.0, height=10.0, ground_w=14.0)
    for specs, fused in ((at, True), (at[:1] + [over, small], False)):
        print(f"size boundary: {[sum(1 for op in s if op[0] in ('disc', 'rect')) for s in specs]} bodies, "
              f"fused expected {fused}", flush=True)
        gw = _worlds(gpu_ctx, specs)
        gpu_ctx.set_profiling(True)
        gpu_ctx.reset_phase_times()
        try:
            gpu_ctx.process_batch_handles([g.handle for g in gw], DT, 4, 4, 20)
            launches = gpu_ctx.fused_times()[1]
        finally:
            gpu_ctx.set_profiling(False)
        assert (launches > 0) == fused, (fused, launches)
        for g, ops in zip(gw, specs):
            assert g.bodies().shape[0] == sum(1 for op in ops if op[0] in ("disc", "rect"))
            o = _oracle_run(oracle, ops, 20)
            assert bits_equal(g.bodies(), o.bodies()), first_diff(g.bodies(), o.bodies())
            assert bits_equal(g.manifolds(), o.manifolds())
            o.close()
            g.close()
