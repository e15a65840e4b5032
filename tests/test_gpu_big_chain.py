"""Worlds beyond LDS whose touched bodies exceed the full compact state (28 B
each) but fit the light one (momentum and counter in LDS, 16 B each; each
item's inverse masses and inertias carried in its lane slot): bit for bit
against the oracle, cold and warm started.  The world is a dense pile of
7,000 bodies, so nearly every body is touched by a contact from the first
step; the LDS limit of 64 bodies sends it to the grid-wide body / item
kernels and the one-workgroup chain."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff
from resolve2d_amd import scenes
from resolve2d_amd.engine import GpuWorld, apply_scene

pytestmark = pytest.mark.gpu

DT = scenes.DT
N = 7000
# the chain kernel's LDS (160 KB less 1 KB), the full compact state's and the
# light state's bytes per touched body, and the light chain's lane slots
# (256 lanes x 6 rows x 16 B)
LDS, FULL_B, LITE_B, LITE_SLOTS = 160 * 1024 - 1024, 28, 16, 256 * 6 * 16


def _touched(o, ops):
    """Bodies the step's solver items touch (both bodies of every manifold,
    the bodies of every joint), as the chain numbers them."""
    m = o.manifolds()
    ids = set(m[:, 0].astype(int).tolist()) | set(m[:, 1].astype(int).tolist())
    for op in ops:
        if op[0] == "joint":
            ids.add(op[5])
            if op[1] <= 1:
                ids.add(op[6])
    return len(ids)


@pytest.mark.parametrize("warm", [0.0, 0.8], ids=["cold", "warm"])
def test_light_compact_chain_matches_oracle(oracle, warm):
    from resolve2d_amd.solver import Context
    ctx = Context(0)
    ctx.set_lds_limit(64)
    ops = scenes.mixed(N, seed=5, width=60.0, height=60.0, ground_w=80.0)
    g, o = GpuWorld(ctx), oracle.OracleWorld("f32")
    for e in (g, o):
        apply_scene(ops, e)
        if warm:
            e.set_warm_start(warm)
    lite_steps = 0
    for step in range(12):
        print(f"light compact chain ({'warm' if warm else 'cold'}): step {step + 1} of 12", flush=True)
        assert g.process(DT, 4, 4) == 0
        assert o.process(DT, 4, 4) == 0
        nc = _touched(o, ops)
        if nc * FULL_B > LDS and nc * LITE_B + LITE_SLOTS <= LDS:
            lite_steps += 1
        a, b = g.bodies(), o.bodies()
        assert bits_equal(a, b), (step, first_diff(a, b))
    mg, mo = g.manifolds(), o.manifolds()
    assert mg.shape == mo.shape and bits_equal(mg, mo)
    # the case this test is for: the light state carried the chain
    assert lite_steps >= 6, lite_steps
    g.close()
    o.close()
    ctx.close()
