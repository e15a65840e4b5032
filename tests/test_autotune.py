"""Autotuner host logic (no GPU): shape keys, the median pick, the cache."""
import os

from resolve2d_amd import autotune


def test_pow2_shape_key():
    assert autotune._pow2(1) == 1 and autotune._pow2(5) == 8 and autotune._pow2(1024) == 1024
    assert autotune.shape_key((1024, 512, 262144)) == "w1024_n512_b262144"


def test_pick_median_and_ties():
    t = {(1, 0): [3.0, 1.0, 2.0], (2, 0): [2.0, 2.0, 9.0], (3, 0): [1.5, 5.0, 1.9]}
    assert autotune.pick(t) == (3, 0)        # medians 2.0, 2.0, 1.9
    t = {(4, 0): [1.0], (5, 0): [1.0]}
    assert autotune.pick(t) == (4, 0)        # ties keep the earlier candidate


def test_candidates_cover_variants():
    c = autotune.candidates(lds_limits=(0, 1024))
    # phase kernels: every variant x both LDS limits; fused: the variants with a fused kernel
    assert len(c) == 2 * len(autotune.PHASE_VARIANTS) + len(autotune.FUSED_VARIANTS)
    assert (1, 0, False) in c and (7, 1024, False) in c and (3, 0, True) in c
    assert (0, 0, True) not in c and (8, 0, True) not in c and (3, 1024, True) not in c
    assert autotune.candidates(variants=[3], lds_limits=(0,), fused=(True,)) == [(3, 0, True)]


def test_cache_roundtrip(tmp_path):
    p = os.path.join(str(tmp_path), "sub", "tune.json")
    assert autotune.load_cache(p) == {}
    autotune.save_cache({"w8_n512_b2048": {"variant": 3, "lds_limit": 0}}, p)
    assert autotune.load_cache(p)["w8_n512_b2048"]["variant"] == 3
    with open(p, "w") as f:
        f.write("not json")
    assert autotune.load_cache(p) == {}
