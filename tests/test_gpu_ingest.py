"""Bulk scene ingestion (r2d_make_bodies; EntityFactory.appendBody order,
lib.zig:66-71): from 32 bodies on the device builds the bodies straight into
the world's resident state.  Checked bit for bit against the same bodies made
one call at a time, and against the oracle after removals, appends to a
stepped world (the device copy is the current one) and growth of the world's
device block, over steps."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _bodies(rng, n):
    """Mixed discs and rectangles with every option used: velocities,
    angles, angular velocities, density or mass, friction."""
    kind = rng.integers(0, 2, n).astype(np.int32)
    f = lambda lo, hi: rng.uniform(lo, hi, n).astype(np.float32)
    return dict(kind=kind, px=f(-30, 30), py=f(1, 60), vx=f(-2, 2), vy=f(-2, 2), angle=f(0, 6.3),
                omega=f(-3, 3), mu=f(0.1, 0.9), mass_kind=rng.integers(0, 2, n).astype(np.int32),
                mass_value=f(0.5, 3.0), dim_a=f(0.2, 1.0), dim_b=f(0.3, 1.0))


def _bulk(w, b):
    return w.make_bodies(b["kind"], b["px"], b["py"], b["vx"], b["vy"], b["angle"], b["omega"], b["mu"],
                         b["mass_kind"], b["mass_value"], b["dim_a"], b["dim_b"])


def _single(e, b):
    ids = []
    for i in range(len(b["kind"])):
        common = (float(b["px"][i]), float(b["py"][i]), float(b["vx"][i]), float(b["vy"][i]),
                  float(b["angle"][i]), float(b["omega"][i]), float(b["mu"][i]), int(b["mass_kind"][i]),
                  float(b["mass_value"][i]))
        if b["kind"][i] == 0:
            ids.append(e.make_disc(*common, float(b["dim_a"][i])))
        else:
            ids.append(e.make_rect(*common, float(b["dim_a"][i]), float(b["dim_b"][i])))
    return ids


def test_bulk_equals_single_calls(gpu_ctx):
    from resolve2d_amd.engine import GpuWorld
    rng = np.random.default_rng(3)
    for n in (31, 32, 500):       # the host path, the first device size, a larger one
        b = _bodies(rng, n)
        bulk, single = GpuWorld(gpu_ctx), GpuWorld(gpu_ctx)
        ids = _bulk(bulk, b)
        ids1 = _single(single, b)
        assert list(ids) == ids1 == list(range(n))
        gb, gs = bulk.bodies(), single.bodies()
        assert bits_equal(gb, gs), (n, first_diff(gb, gs))
        bulk.close()
        single.close()


def test_appends_to_a_stepped_world_match_the_oracle(gpu_ctx, oracle):
    """Ground + 20 bodies (single calls), 10 steps, a removal (swapRemove),
    then 300 bodies in one bulk call -- past the device block's capacity --
    10 more steps, 40 bodies more in a second bulk call, 20 steps: the state
    equals the oracle's doing the same with single calls."""
    from resolve2d_amd.engine import GpuWorld
    rng = np.random.default_rng(7)
    g, o = GpuWorld(gpu_ctx), oracle.OracleWorld("f32")
    for e in (g, o):
        e.make_gravity(float(np.float32(9.81)))
        gnd = e.make_rect(0.0, -0.5, 0.0, 0.0, 0.0, 0.0, float(np.float32(0.4)), 0, 1.0, 100.0, 1.0)
        e.set_static(gnd, True)
    first = _bodies(rng, 20)
    _single(g, first)
    _single(o, first)
    for _ in range(10):
        assert g.process(DT, 4, 4) == 0 and o.process(DT, 4, 4) == 0
    for e in (g, o):
        e.remove_body(7)
    big = _bodies(rng, 300)
    ids = _bulk(g, big)
    ids_o = _single(o, big)
    assert list(ids) == ids_o
    assert bits_equal(g.bodies(), o.bodies()), first_diff(g.bodies(), o.bodies())
    for _ in range(10):
        assert g.process(DT, 4, 4) == 0 and o.process(DT, 4, 4) == 0
    more = _bodies(rng, 40)
    assert list(_bulk(g, more)) == _single(o, more)
    for _ in range(20):
        assert g.process(DT, 4, 4) == 0 and o.process(DT, 4, 4) == 0
    gb, ob = g.bodies(), o.bodies()
    assert bits_equal(gb, ob), first_diff(gb, ob)
    assert bits_equal(g.manifolds(), o.manifolds())
    g.close()
