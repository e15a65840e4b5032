"""Warm starting in the oracle (an extension: the reference rebuilds every
manifold with zero impulses, lib.zig:261 and collision.zig:38-51; its
roadmap asks for manifolds kept across frames, README.md:60-61).

No reference output covers it, so each rule is pinned by what physics or the
rule itself fixes: off and the first step are the reference's bits; the
carried impulses alone conserve linear and angular momentum (a wrong sign,
lever arm or body fails); the table keeps at most 8 manifolds per lower
body; a turned normal, a removal or a new factor starts cold; and a resting
pyramid is steadier with it than without."""
import math

import numpy as np

from resolve2d_amd import scenes
from resolve2d_amd.scenes import _Builder
from resolve2d_amd.engine import apply_scene

DT = 1.0 / 60.0


def world(oracle, ops, prec="f32", warm=0.0):
    w = oracle.OracleWorld(prec)
    apply_scene(ops, w)
    if warm:
        w.set_warm_start(warm)
    return w


def bits(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).view(np.uint32)


def test_off_and_first_step_are_the_reference(oracle):
    ops = scenes.pyramid(rows=8)
    cold, off, warm = world(oracle, ops), world(oracle, ops), world(oracle, ops, warm=0.8)
    off.set_warm_start(0.0)
    for w in (cold, off, warm):
        assert w.process(DT, 4, 4) == 0
    # nothing is kept before the first step: identical bits
    assert (bits(warm.bodies()) == bits(cold.bodies())).all() and warm.num_warm() == 0
    for _ in range(20):
        for w in (cold, off, warm):
            assert w.process(DT, 4, 4) == 0
    assert (bits(off.bodies()) == bits(cold.bodies())).all()
    assert off.num_warm() == 0
    assert warm.num_warm() > 0 and not (bits(warm.bodies()) == bits(cold.bodies())).all()


def _cluster():
    """Two boxes and a disc closing on each other in free space (no gravity,
    no static body): every contact impulse acts between two dynamic bodies,
    and the first step's collision leaves impulses to carry."""
    b = _Builder()
    b.rect(0.0, 0.0, 1.0, 1.0, angle=0.05, vx=1.0, omega=0.2)
    b.rect(0.999, 0.1, 1.0, 1.0, angle=-0.03, vx=-1.0)
    b.disc(0.2, 0.999, 0.5, vy=-1.5)
    return b.ops


def _momenta(w):
    b = w.bodies()
    p = b[:, 5:7].sum(axis=0)
    lz = float(np.sum(b[:, 11] + b[:, 3] * b[:, 6] - b[:, 4] * b[:, 5]))
    return p, lz


def test_carried_impulses_conserve_momentum(oracle):
    """One step of one sub-step and no iterations applies only the carried
    impulses, before any position changes: each acts at the same point of
    both bodies (ref.pos + r1 = inc.pos + r2), so the linear and the angular
    momentum about the origin stay put (f64), while single bodies change."""
    w = world(oracle, _cluster(), "f64", warm=1.0)
    assert w.process(DT, 4, 4) == 0 and len(w.manifolds()) == 3
    before = w.bodies()
    p0, l0 = _momenta(w)
    assert w.process(DT, 1, 0) == 0
    assert w.num_warm() == 3
    after = w.bodies()
    p1, l1 = _momenta(w)
    moved = np.abs(after[:, 5:7] - before[:, 5:7]).max()
    assert moved > 1e-3, "no carried impulse was applied"
    assert np.abs(p1 - p0).max() <= 1e-12 * max(1.0, np.abs(before[:, 5:7]).sum())
    assert abs(l1 - l0) <= 1e-12 * max(1.0, abs(l0) + moved)
    # a new factor empties the kept manifolds
    w.set_warm_start(0.5)
    assert w.process(DT, 1, 0) == 0 and w.num_warm() == 0


def test_factor_scales_the_carried_impulse(oracle):
    """Two worlds with the same history at factor 1; the second then carries
    at factor 0.5 (kept manifolds unchanged): on a step of one sub-step and
    no iterations every body's momentum change is half the first world's."""
    ws = []
    for f in (1.0, 0.5):
        w = world(oracle, _cluster(), "f64", warm=1.0)
        assert w.process(DT, 4, 4) == 0
        w.set_warm_factor(f)
        before = w.bodies()
        assert w.process(DT, 1, 0) == 0
        ws.append(w.bodies()[:, 5:7] - before[:, 5:7])
    assert np.abs(ws[0]).max() > 1e-3
    np.testing.assert_allclose(ws[1], 0.5 * ws[0], rtol=1e-9, atol=1e-12)


def test_table_keeps_eight_per_lower_body(oracle):
    """Twelve boxes resting apart on one static ground: with the ground in slot
    0 every manifold is kept under slot 0, of which 8 survive; with the ground
    last each box is its own manifold's lower body and all 12 survive."""
    for ground_first, want in ((True, 8), (False, 12)):
        b = _Builder()
        b.gravity(9.81)
        if ground_first:
            b.static(b.rect(0.0, -0.5, 60.0, 1.0))
        for i in range(12):
            b.rect(-22.0 + 4.0 * i, 0.499, 1.0, 1.0)
        if not ground_first:
            b.static(b.rect(0.0, -0.5, 60.0, 1.0))
        w = world(oracle, b.ops, warm=0.8)
        for _ in range(3):
            assert w.process(DT, 4, 4) == 0
        assert len(w.manifolds()) == 12
        assert w.num_warm() == want


def test_turned_normal_and_removal_start_cold(oracle):
    """Two touching discs: moved so that the contact normal turns by 2 degrees
    the manifold is carried; by 90 degrees it is not (cos < 0.99).  Removing
    any body empties the kept manifolds."""
    def pair():
        b = _Builder()
        b.disc(0.0, 0.0, 0.5)
        b.disc(0.999, 0.0, 0.5)
        b.disc(10.0, 10.0, 0.3)
        return b.ops

    for turn, want in ((2.0, 1), (90.0, 0)):
        w = world(oracle, pair(), warm=1.0)
        assert w.process(DT, 4, 4) == 0 and len(w.manifolds()) == 1
        a = math.radians(turn)
        w.set_body(1, 0.999 * math.cos(a), 0.999 * math.sin(a), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
        assert w.process(DT, 4, 4) == 0 and len(w.manifolds()) == 1
        assert w.num_warm() == want, turn
    w = world(oracle, pair(), warm=1.0)
    for _ in range(2):
        assert w.process(DT, 4, 4) == 0
    assert w.num_warm() == 1
    w.remove_body(2)
    assert w.process(DT, 4, 4) == 0 and w.num_warm() == 0


def test_resting_pyramid_is_steadier(oracle):
    """A 10-row pyramid of unit boxes at rest, 200 steps: with warm starting
    (factor 0.8) the deepest contact and the energy drift are smaller."""
    def energy(b):
        dyn = b[:, 2] == 0
        m, inertia = b[dyn, 9], b[dyn, 13]
        return float(np.sum(0.5 * (b[dyn, 5] ** 2 + b[dyn, 6] ** 2) / m + 0.5 * b[dyn, 11] ** 2 / inertia
                            + m * np.float32(9.81) * b[dyn, 4]))

    def max_penetration(man):
        return max([0.0] + [float(-man[man[:, 7 + 13 * p] != 0, 16 + 13 * p].min())
                            for p in range(2) if (man[:, 7 + 13 * p] != 0).any()])

    out = {}
    for f in (0.0, 0.8):
        w = world(oracle, scenes.pyramid(rows=10), warm=f)
        e0 = energy(w.bodies())
        pen = 0.0
        for s in range(200):
            assert w.process(DT, 4, 4) == 0
            if s >= 100:
                pen = max(pen, max_penetration(w.manifolds()))
        out[f] = (pen, abs(energy(w.bodies()) - e0) / abs(e0))
    assert out[0.8][0] < out[0.0][0] and out[0.8][1] < out[0.0][1], out
