"""The oracle's f32 sin/cos (the msun/musl method of Zig's compiler_rt, used by
rotate2, nmath.zig:129-136) against FP64 sin/cos and known values."""
import math
import struct

import numpy as np
import pytest


def ulp_err(f, exact):
    f = np.float32(f)
    if exact == 0.0:
        return 0.0 if f == 0.0 else float("inf")
    up = np.nextafter(np.float32(abs(exact)), np.float32(np.inf)) - np.float32(abs(exact))
    return abs(float(f) - exact) / float(up)


def samples(seed=0):
    rng = np.random.default_rng(seed)
    xs = np.concatenate([
        rng.uniform(-1, 1, 3000), rng.uniform(-8, 8, 4000), rng.uniform(-1e3, 1e3, 2000),
        rng.uniform(-1e6, 1e6, 500), rng.uniform(-4e8, 4e8, 300), rng.uniform(-1e20, 1e20, 100),
        rng.uniform(-1e-3, 1e-3, 200),
    ]).astype(np.float32)
    # every branch threshold of sinf/cosf and its neighbours
    th = [0x3f490fda, 0x39800000, 0x407b53d1, 0x4016cbe3, 0x40e231d5, 0x40afeddf, 0x4dc90fdb]
    extra = []
    for t in th:
        for d in (-2, -1, 0, 1, 2):
            v = struct.unpack("<f", struct.pack("<I", t + d))[0]
            extra += [v, -v]
    return np.concatenate([xs, np.array(extra, np.float32)])


def test_sin_cos_within_one_ulp(oracle):
    worst_s = worst_c = 0.0
    cr = 0
    xs = samples()
    for x in xs:
        xd = float(x)
        s, c = oracle.sinf(xd), oracle.cosf(xd)
        es, ec = ulp_err(s, math.sin(xd)), ulp_err(c, math.cos(xd))
        worst_s, worst_c = max(worst_s, es), max(worst_c, ec)
        cr += int(np.float32(s) == np.float32(math.sin(xd))) + int(np.float32(c) == np.float32(math.cos(xd)))
    assert worst_s <= 1.0 and worst_c <= 1.0, (worst_s, worst_c)
    # msun's float kernels round correctly for all but a sliver of inputs
    assert cr / (2 * len(xs)) > 0.995


def test_known_values(oracle):
    assert oracle.sinf(0.0) == 0.0 and oracle.cosf(0.0) == 1.0
    assert math.copysign(1, oracle.sinf(-0.0)) == -1.0
    # |x| < 2^-12 returns x (sin) and 1 (cos)
    assert oracle.sinf(1e-5) == float(np.float32(1e-5)) and oracle.cosf(1e-5) == 1.0
    assert math.isnan(oracle.sinf(float("inf"))) and math.isnan(oracle.cosf(float("nan")))
    pi2 = float(np.float32(math.pi / 2))
    assert oracle.sinf(pi2) == 1.0
    assert abs(oracle.cosf(pi2) - math.cos(pi2)) < 1e-12


def test_periodic_consistency(oracle):
    # sin^2 + cos^2 stays within a few ulp of one across all reduction ranges
    for x in samples(1)[::7]:
        s, c = oracle.sinf(float(x)), oracle.cosf(float(x))
        if math.isfinite(s):
            assert abs(s * s + c * c - 1.0) < 4e-7


def test_huge_argument_reduction(oracle):
    # beyond 2^28*pi/2 the reduction uses a 128-bit window of 2/pi; exact FP64 is the check
    rng = np.random.default_rng(7)
    for x in rng.uniform(4.3e8, 3e38, 400).astype(np.float32):
        xd = float(x)
        assert ulp_err(oracle.sinf(xd), math.sin(xd)) <= 1.0
        assert ulp_err(oracle.cosf(xd), math.cos(xd)) <= 1.0
