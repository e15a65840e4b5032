"""Serving: continuous batching over one context (r2d_serve_*).

Each served world must end bit-identical to the same scene stepped alone
(one world per launch), whatever else shares its batches; admission must
respect the HBM budget and arrival order; worlds held by the server refuse
edits.  The mix covers every solve route: worlds small enough for the
register-slot kernel, worlds whose body state fits LDS, and a world too large
for LDS (body state in global memory)."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff

pytestmark = pytest.mark.gpu

STEPS = 24


def _served_world(ctx, seed, n):
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld
    return scenes.build_served(GpuWorld(ctx), seed, n)


def _scene_world(ctx, ops):
    from resolve2d_amd.engine import GpuWorld, apply_scene
    w = GpuWorld(ctx)
    apply_scene(ops, w)
    return w


def _mix(ctx):
    from resolve2d_amd import scenes
    specs = [("served", 11, 16), ("served", 12, 700), ("ops", scenes.car_demo()), ("served", 13, 150),
             ("served", 14, 6100), ("ops", scenes.batched_world(31)), ("served", 15, 1500), ("served", 16, 40),
             ("ops", scenes.pyramid(rows=12))]
    out = []
    for sp in specs:
        out.append((sp, _served_world(ctx, sp[1], sp[2]) if sp[0] == "served" else _scene_world(ctx, sp[1])))
    return out


def _alone(ctx, sp):
    w = _served_world(ctx, sp[1], sp[2]) if sp[0] == "served" else _scene_world(ctx, sp[1])
    for _ in range(STEPS):
        assert w.process(1.0 / 60.0, 4, 4) == 0
    return w


@pytest.mark.parametrize("per_tick", [1, 5])
def test_served_results_equal_alone(gpu_ctx, per_tick):
    """per_tick: steps a lane's launch takes (5: fewer when a running world
    has fewer left, so requests still end after exactly STEPS steps)."""
    from resolve2d_amd.serve import Server
    mix = _mix(gpu_ctx)
    fps = [w.footprint() for _, w in mix]
    # room for about three of the larger worlds at once: requests queue and
    # are admitted as others finish, at every stage of their neighbours' runs
    budget = sorted(fps)[-1] + sorted(fps)[-2] + sorted(fps)[-3]
    sv = Server(gpu_ctx, budget, max_worlds=4, steps_per_tick=per_tick)
    try:
        tickets = []
        order = list(range(len(mix)))
        # three arrive at once, the rest one every third tick
        for i in order[:3]:
            tickets.append(sv.submit(mix[i][1], STEPS))
        nxt = 3
        ticks = 0
        finished = []
        while True:
            info = sv.tick()
            ticks += 1
            done_now = sv.completed()
            assert len(done_now) == info["completed"]
            finished += done_now
            assert info["resident_bytes"] <= budget
            assert info["running"] <= 4
            if nxt < len(mix) and ticks % 3 == 0:
                tickets.append(sv.submit(mix[order[nxt]][1], STEPS))
                nxt += 1
            st = sv.stats()
            if nxt == len(mix) and st["completed"] == len(mix) and st["running"] == 0:
                break
            assert ticks < 10 * STEPS * len(mix)
        st = sv.stats()
        assert st["peak_resident_bytes"] <= budget
        assert st["scene_steps"] == STEPS * len(mix)
        assert st["body_steps"] == STEPS * sum(w.bodies().shape[0] for _, w in mix)
        assert st["p99_step_ms"] >= st["p50_step_ms"] > 0
        assert st["lanes"] == 3
        lanes = set()
        for t in tickets:
            p = sv.poll(t)
            assert p["state"] == 2 and p["status"] == 0 and p["steps_done"] == STEPS, p
            lanes.add(p["lane"])
        assert lanes == {0, 1, 2}
        # every request reported finished exactly once
        assert sorted(finished) == sorted(tickets) and sv.completed() == []
    finally:
        sv.close()
    for i, (sp, w) in enumerate(mix):
        ref = _alone(gpu_ctx, sp)
        a, b = w.bodies(), ref.bodies()
        assert bits_equal(a, b), (i, sp[:1] + sp[1:3] if sp[0] == "served" else sp[0], first_diff(a, b))
        w.close()
        ref.close()


def test_admission_fifo_budget_and_busy(gpu_ctx):
    from resolve2d_amd import _native as N
    from resolve2d_amd.serve import Server
    ws = [_served_world(gpu_ctx, 40 + i, n) for i, n in enumerate([300, 50, 300, 20])]
    fps = [w.footprint() for w in ws]
    budget = fps[0] + fps[1]            # the first two fit together, the third must wait
    sv = Server(gpu_ctx, budget, max_worlds=8)
    try:
        big = _served_world(gpu_ctx, 99, 2000)
        with pytest.raises(N.R2DError) as e:
            sv.submit(big, 5)           # can never fit
        assert e.value.status == 8
        big.close()
        t = [sv.submit(w, 3) for w in ws]
        # a world held by the server refuses edits and direct steps
        assert gpu_ctx.lib.r2d_make_gravity(ws[0].handle, 1.0) == 9
        assert ws[0].process(1.0 / 60.0, 4, 4) == 9
        info = sv.tick()
        assert info["admitted"] == 2 and info["queued"] == 2
        # FIFO: the small fourth request does not overtake the third
        assert sv.poll(t[3])["state"] == 0 and sv.poll(t[2])["state"] == 0
        sv.drain()
        for x in t:
            assert sv.poll(x)["state"] == 2
        assert sv.stats()["peak_resident_bytes"] <= budget
        # released: editable again
        assert gpu_ctx.lib.r2d_make_gravity(ws[0].handle, 0.0) == 0
    finally:
        sv.close()
        for w in ws:
            w.close()


def test_global_memory_route_matches_oracle(gpu_ctx, oracle):
    """A world too large for LDS body state (6,100 bodies) against the oracle."""
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import apply_scene
    ops = scenes.served_ops(7, 6100)
    g = _served_world(gpu_ctx, 7, 6100)
    o = oracle.OracleWorld("f32")
    apply_scene(ops, o)
    assert bits_equal(g.bodies(), o.bodies())
    for s in range(6):
        assert g.process(1.0 / 60.0, 4, 4) == 0
        assert o.process(1.0 / 60.0, 4, 4) == 0
    a, b = g.bodies(), o.bodies()
    assert bits_equal(a, b), first_diff(a, b)
    assert bits_equal(g.manifolds(), o.manifolds())
    g.close()


def test_finished_worlds_may_be_destroyed_before_the_server(gpu_ctx):
    """A finished request releases its world: the caller may free it while the
    server lives on, and closing the server touches only unfinished ones."""
    from resolve2d_amd.serve import Server
    sv = Server(gpu_ctx, 1 << 30, max_worlds=8)
    ws = [_served_world(gpu_ctx, 60 + i, 30 + 10 * i) for i in range(4)]
    t = [sv.submit(w, 2) for w in ws[:3]]
    sv.drain()
    for w in ws[:3]:
        w.close()
    late = sv.submit(ws[3], 50)
    sv.tick()
    sv.close()                      # releases the unfinished request's world
    assert sv.handle is None
    assert ws[3].process(1.0 / 60.0, 4, 4) == 0
    assert all(x > 0 for x in t + [late])
    ws[3].close()


def test_served_warm_worlds_equal_alone(gpu_ctx):
    """Warm-started worlds (r2d_set_warm_start) served beside cold ones: the
    lanes' batches take the warm kernels, and every world ends bit-identical
    to the same scene stepped alone with the same factor."""
    from resolve2d_amd import scenes
    from resolve2d_amd.serve import Server
    specs = [("served", 21, 300), ("ops", scenes.pyramid(rows=10)), ("served", 22, 2000),
             ("ops", scenes.batched_world(41)), ("served", 23, 90)]
    warms = [0.8, 1.0, 0.8, 0.0, 0.5]
    mix = []
    for sp, f in zip(specs, warms):
        w = _served_world(gpu_ctx, sp[1], sp[2]) if sp[0] == "served" else _scene_world(gpu_ctx, sp[1])
        w.set_warm_start(f)
        mix.append(w)
    sv = Server(gpu_ctx, sum(w.footprint() for w in mix), max_worlds=3, steps_per_tick=3)
    try:
        tickets = [sv.submit(w, STEPS) for w in mix]
        for _ in range(10 * STEPS * len(mix)):
            sv.tick()
            if sv.stats()["completed"] == len(mix):
                break
        for t in tickets:
            p = sv.poll(t)
            assert p["state"] == 2 and p["status"] == 0 and p["steps_done"] == STEPS, p
    finally:
        sv.close()
    for i, (sp, f, w) in enumerate(zip(specs, warms, mix)):
        ref = _served_world(gpu_ctx, sp[1], sp[2]) if sp[0] == "served" else _scene_world(gpu_ctx, sp[1])
        ref.set_warm_start(f)
        for _ in range(STEPS):
            assert ref.process(1.0 / 60.0, 4, 4) == 0
        a, b = w.bodies(), ref.bodies()
        assert bits_equal(a, b), (i, f, first_diff(a, b))
        w.close()
        ref.close()
