"""GPU step vs the oracle at the benchmark configurations' real sizes, and
the edge cases of the reference's semantics (empty worlds, swapRemove,
setters between steps, static toggling, zero sub-steps / iterations, invalid
constraint ids, non-finite state, the tuned hash mode)."""
import math

import numpy as np
import pytest

from helpers import bits_equal, first_diff, run_pair
from resolve2d_amd import scenes
from resolve2d_amd.engine import GpuWorld, apply_scene
from resolve2d_amd.scenes import _Builder

pytestmark = pytest.mark.gpu


def pair(gpu_ctx, oracle, scene):
    g, o = GpuWorld(gpu_ctx), oracle.OracleWorld("f32")
    apply_scene(scene, g)
    apply_scene(scene, o)
    return g, o


def step_both(g, o, n, dt=scenes.DT, sub=4, iters=4):
    for _ in range(n):
        rg, ro = g.process(dt, sub, iters), o.process(dt, sub, iters)
        assert rg == 0 and ro == 0


def assert_same(g, o, what=""):
    a, b = g.bodies(), o.bodies()
    assert bits_equal(a, b), (what, first_diff(a, b))


def test_config1_pyramid_full_horizon(gpu_ctx, oracle):
    """20-row box pyramid (210 boxes + ground), 600 steps, bit-exact."""
    step, diff = run_pair(scenes.pyramid(rows=20), GpuWorld(gpu_ctx), oracle.OracleWorld("f32"), 600,
                          check_every=100)
    assert step is None, (step, diff)


def test_config2_mixed_4096_full_horizon(gpu_ctx, oracle):
    """4,096 mixed bodies in a 100 x 200 m box, 300 steps, bit-exact."""
    g, o = pair(gpu_ctx, oracle, scenes.mixed(4096, seed=1))
    for k in range(3):
        step_both(g, o, 100)
        assert_same(g, o, f"after {100 * (k + 1)} steps")
    mg, mo = g.manifolds(), o.manifolds()
    assert mg.shape[0] > 1000 and bits_equal(mg, mo)


def test_config2_tuned_hash_full_horizon(gpu_ctx, oracle):
    """Configuration 2 as stated: 4,096 mixed bodies with the tuned spatial
    hash (cell = twice the largest body extent, 2N buckets), 300 steps,
    bit-exact in the reference order; and its coloured schedule against the
    oracle's coloured schedule."""
    sc = scenes.mixed(4096, seed=1, hash_cell="auto")
    assert sc[0][0] == "hash" and 2.5 < sc[0][1] < 2.9
    g, o = pair(gpu_ctx, oracle, sc)
    for k in range(3):
        step_both(g, o, 100)
        assert_same(g, o, f"after {100 * (k + 1)} steps")
    mg, mo = g.manifolds(), o.manifolds()
    assert mg.shape[0] > 1000 and bits_equal(mg, mo)
    gc, oc = pair(gpu_ctx, oracle, sc)
    gc.set_schedule(1)
    oc.set_schedule(1)
    step_both(gc, oc, 60)
    assert_same(gc, oc, "coloured, 60 steps")
    for w in (g, gc):
        w.close()


def test_config3_batched_sample(gpu_ctx, oracle):
    """The benchmark batch (1,024 worlds, U[50,500] bodies, joints, springs)
    stepped together for 240 steps; eight worlds spread over the batch are
    stepped alone on the oracle and must match bit for bit."""
    worlds, specs = [], []
    for i in range(1024):
        sc = scenes.batched_world(1000 + i)
        w = GpuWorld(gpu_ctx)
        apply_scene(sc, w)
        worlds.append(w)
        specs.append(sc)
    gpu_ctx.process_batch_handles([w.handle for w in worlds], scenes.DT, 4, 4, 240)
    for i in (0, 101, 257, 511, 640, 777, 900, 1023):
        o = oracle.OracleWorld("f32")
        apply_scene(specs[i], o)
        for _ in range(240):
            assert o.process(scenes.DT, 4, 4) == 0
        a, b = worlds[i].bodies(), o.bodies()
        assert bits_equal(a, b), (i, first_diff(a, b))
    for w in worlds:
        w.close()


def test_config3_every_world_bit_exact(gpu_ctx, oracle):
    """The whole benchmark batch (1,024 worlds, 286,693 bodies) after its
    240-step horizon in one call: every world bit for bit against the oracle
    stepping it alone (the oracle runs in host processes)."""
    import os
    from oracle_pool import oracle_bodies
    specs = [scenes.batched_world(1000 + i) for i in range(1024)]
    worlds = []
    for sc in specs:
        w = GpuWorld(gpu_ctx)
        apply_scene(sc, w)
        worlds.append(w)
    print("config 3, every world: 1,024 worlds x 240 steps on the GPU", flush=True)
    gpu_ctx.process_batch_handles([w.handle for w in worlds], scenes.DT, 4, 4, 240)
    procs = max(1, min(15, (os.cpu_count() or 2) - 1))
    print(f"config 3, every world: the oracle over {procs} host processes", flush=True)
    ref = oracle_bodies(specs, 240, scenes.DT, 4, 4, procs)
    bad = [i for i, (w, b) in enumerate(zip(worlds, ref)) if not bits_equal(w.bodies(), b)]
    assert not bad, (len(bad), bad[:8], first_diff(worlds[bad[0]].bodies(), ref[bad[0]]))
    for w in worlds:
        w.close()


def test_empty_and_static_only_worlds(gpu_ctx, oracle):
    g = GpuWorld(gpu_ctx)
    assert g.process(scenes.DT, 4, 4) == 0 and g.bodies().shape == (0, 19)
    b = _Builder()
    b.gravity(9.81)
    for i in range(5):
        b.static(b.rect(i * 0.9, 0.0, 1.0, 1.0))
    g2, o2 = pair(gpu_ctx, oracle, b.ops)
    step_both(g2, o2, 3)
    assert_same(g2, o2)
    assert g2.manifolds().shape[0] == 0


def test_swap_remove_and_setters_between_steps(gpu_ctx, oracle):
    g, o = pair(gpu_ctx, oracle, scenes.mixed(200, seed=9, width=20.0, height=20.0, ground_w=30.0))
    step_both(g, o, 20)
    for e in (g, o):
        e.remove_body(5)
        e.remove_body(77)
        e.set_body(10, 1.0, 5.0, 3.0, -2.0, 0.5, 40.0, 0.3, 1.5, 2.0)
        e.set_static(12, True)
        e.set_static(0, False)   # the ground starts falling
    step_both(g, o, 20)
    assert_same(g, o, "after removals and setters")
    ids = g.bodies()[:, 0]
    assert 5 not in ids and 77 not in ids and len(ids) == 199


def test_zero_substeps_and_iterations(gpu_ctx, oracle):
    g, o = pair(gpu_ctx, oracle, scenes.pyramid(rows=4))
    for sub, it in ((0, 4), (4, 0), (1, 1), (3, 2)):
        step_both(g, o, 2, sub=sub, iters=it)
        assert_same(g, o, (sub, it))
        # after a step without sub-steps no preStep ran: tangent, friction and
        # effective masses read back as their initial zeros
        assert bits_equal(g.manifolds(), o.manifolds()), (sub, it)


def test_invalid_constraint_id_is_an_error(gpu_ctx, oracle):
    from resolve2d_amd._native import R2DError
    b = _Builder()
    b.gravity(9.81)
    a = b.disc(0.0, 0.0, 0.3)
    c = b.disc(1.0, 0.0, 0.3)
    b.joint(0, a, c, target=1.0)
    g, o = pair(gpu_ctx, oracle, b.ops)
    g.remove_body(c)
    o.remove_body(c)
    assert g.process(scenes.DT, 4, 4) == 2          # R2D_E_INVALID_ID
    assert o.process(scenes.DT, 4, 4) == -2         # error.InvalidRigidBodyId
    assert g.process(scenes.DT, 0, 4) == 0          # never solved without sub-steps


def test_nonfinite_state_is_reported_not_stepped(gpu_ctx):
    b = _Builder()
    b.gravity(9.81)
    b.disc(0.0, 0.0, 0.3)
    b.disc(2.0, 0.0, 0.3)
    g = GpuWorld(gpu_ctx)
    apply_scene(b.ops, g)
    g.process(scenes.DT, 4, 4)
    g.set_body(1, float("nan"), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    g.process(scenes.DT, 4, 4)      # AABB refreshed with a NaN centre
    before = g.bodies()
    assert g.process(scenes.DT, 4, 4) == 4        # R2D_E_NONFINITE
    assert bits_equal(before, g.bodies())


def test_tuned_hash_mode_matches_oracle(gpu_ctx, oracle):
    """Cell width 2 x the largest extent, 2N buckets (the configuration-2 tuning)."""
    sc = scenes.mixed(600, seed=4, width=30.0, height=40.0, ground_w=40.0, hash_cell=2.0)
    g, o = pair(gpu_ctx, oracle, sc)
    step_both(g, o, 60)
    assert_same(g, o)


def test_excluded_pairs_and_joint_on_static(gpu_ctx, oracle):
    b = _Builder()
    b.gravity(9.81)
    gnd = b.rect(0.0, -0.5, 30.0, 1.0)
    b.static(gnd)
    prev = b.rect(0.0, 6.0, 0.5, 0.5)
    b.joint(0, gnd, prev, target=6.5)           # a static body written by a joint
    for i in range(1, 12):
        cur = b.rect(0.45 * i, 6.0 - 0.1 * i, 0.5, 0.5)
        b.joint(1, prev, cur, r1=(0.2, 0.0), r2=(-0.2, 0.0), target=0.1)
        b.exclude(prev, cur)
        prev = cur
    b.joint(3, prev, target=2.0, pmax=5.0, pmin=-5.0)
    b.joint(2, prev, tpos=(3.0, 3.0), beta=3.0)
    g, o = pair(gpu_ctx, oracle, b.ops)
    step_both(g, o, 120)
    assert_same(g, o)
