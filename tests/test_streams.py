"""Served-scene generators and the arrival stream (host logic, no GPU)."""
import math

import numpy as np

from resolve2d_amd import scenes


def test_served_scene_is_seeded_and_shaped():
    a = scenes.served_arrays(3, 500)
    b = scenes.served_arrays(3, 500)
    for k in ("kind", "px", "py", "angle", "dim_a", "dim_b"):
        assert np.array_equal(a[k], b[k])
        assert a[k].shape == (500,)
    disc = a["kind"] == 0
    assert disc.sum() == 250
    assert (a["dim_a"][disc] >= 0.2).all() and (a["dim_a"][disc] <= 0.6).all()
    assert (a["dim_a"][~disc] >= 0.3).all() and (a["dim_a"][~disc] <= 1.0).all()
    assert (a["dim_b"][~disc] >= 0.3).all() and (a["dim_b"][~disc] <= 1.0).all()
    assert (a["angle"][disc] == 0).all()
    w = float(a["width"])
    assert (np.abs(a["px"]) <= w / 2).all() and (a["py"] >= 1.0).all() and (a["py"] <= 1.0 + 2 * w).all()


def test_served_ops_match_arrays():
    a = scenes.served_arrays(9, 30)
    ops = scenes.served_ops(9, 30)
    assert ops[0][0] == "gravity" and ops[1][0] == "rect" and ops[2] == ("static", 0, True)
    bodies = ops[3:]
    assert len(bodies) == 30
    for i, op in enumerate(bodies):
        assert op[0] == ("disc" if a["kind"][i] == 0 else "rect")
        assert np.float32(op[1]) == a["px"][i] and np.float32(op[2]) == a["py"][i]


def test_served_ops_run_on_the_oracle(oracle):
    from resolve2d_amd.engine import apply_scene
    o = oracle.OracleWorld("f32")
    apply_scene(scenes.served_ops(4, 60), o)
    for _ in range(3):
        assert o.process(scenes.DT, 4, 4) == 0
    assert o.bodies().shape == (61, 19)


def test_poisson_stream_rate_and_sizes():
    arr = scenes.poisson_stream(5, 200.0, 4000)
    t = np.array([a[0] for a in arr])
    n = np.array([a[1] for a in arr])
    assert (np.diff(t) >= 0).all()
    rate = len(t) / t[-1]
    assert 180 < rate < 220
    assert n.min() >= 16 and n.max() <= 8192
    # log-uniform: the median is the geometric mean of the bounds
    assert abs(math.log(np.median(n)) - 0.5 * (math.log(16) + math.log(8192))) < 0.15


def test_large_constrained_scene_shape():
    from collections import Counter
    ops = scenes.large_constrained()
    c = Counter(o[0] for o in ops)
    assert c["disc"] + c["rect"] == 65540 and c["static"] == 4 and c["gravity"] == 1
    joints = [o for o in ops if o[0] == "joint"]
    dist = [j for j in joints if j[1] == 0]
    mot = [j for j in joints if j[1] == 3]
    assert len(dist) == 2000 and len(mot) == 64
    assert all(j[2] == np.float32(50.0) and j[3] == np.float32(-50.0) for j in mot)
    assert all(-5.0 <= j[11] <= 5.0 for j in mot)
    # ropes: 100 chains of 20 consecutive links
    pairs = sorted((j[5], j[6]) for j in dist)
    assert all(b == a + 1 for a, b in pairs)
    assert len({a for a, _ in pairs} - {b for _, b in pairs}) == 100
