"""The oracle against the reference's own known-answer tests and vectors:
nmath.zig:157-176 ("vector2") and aabb.zig:34-45 ("sanity check aabb
overlapping"), plus hash/constant checks of SpatialHash.zig:78-81."""
import ctypes
import math

import numpy as np


def nm(lib, op, a, b=(0.0, 0.0)):
    out = (ctypes.c_double * 2)()
    lib.or32_nmath.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.POINTER(ctypes.c_double)]
    lib.or32_nmath(op, a[0], a[1], b[0], b[1], out)
    return out[0], out[1]


def test_vector2_kat(oracle):
    lib = oracle.load()
    a, b = (2.0, 3.0), (4.0, 5.0)
    assert nm(lib, 0, a, b) == (6.0, 8.0)          # add2
    assert nm(lib, 1, a, b) == (-2.0, -2.0)        # sub2
    assert nm(lib, 2, a, b) == (8.0, 15.0)         # multelem2
    assert nm(lib, 3, a, b) == (0.5, float(np.float32(0.6)))  # divelem2
    assert nm(lib, 4, a, b)[0] == 23.0             # dot2
    assert nm(lib, 5, a, b)[0] == -2.0             # cross2
    assert nm(lib, 6, a)[0] == 13.0 and nm(lib, 6, b)[0] == 41.0   # length2sq
    assert nm(lib, 7, a)[0] == float(np.sqrt(np.float32(13)))      # length2
    assert nm(lib, 7, b)[0] == float(np.sqrt(np.float32(41)))


def test_nmath_semantics(oracle):
    lib = oracle.load()
    # normalize2: below 1e-3 -> zero vector; otherwise multiply by the reciprocal
    assert nm(lib, 8, (5e-4, 0.0)) == (0.0, 0.0)
    x, y = np.float32(3.0), np.float32(4.0)
    inv = np.float32(1.0) / np.float32(5.0)
    assert nm(lib, 8, (3.0, 4.0)) == (float(x * inv), float(y * inv))
    # rotate90 helpers
    assert nm(lib, 10, (1.0, 2.0)) == (2.0, -1.0)
    assert nm(lib, 11, (1.0, 2.0)) == (-2.0, 1.0)
    # rotate2 by pi/2 (f32) maps x onto y up to the f32 cosine of pi/2
    r = nm(lib, 9, (1.0, 0.0), (float(np.float32(math.pi / 2)), 0.0))
    assert r[1] == 1.0 and abs(r[0]) < 1e-7


def box(x, y, hw, hh):
    return (ctypes.c_double * 4)(x, y, hw, hh)


def test_aabb_kat(oracle):
    lib = oracle.load()
    lib.or32_aabb_intersects.argtypes = [ctypes.POINTER(ctypes.c_double)] * 2
    s1 = box(1, 1, 1.0, 0.5)
    s2 = box(2, 3, 1.5, 3.0)
    assert lib.or32_aabb_intersects(s1, s2)
    s2 = box(2, 10, 1.5, 3.0)
    assert not lib.or32_aabb_intersects(s1, s2)
    s1 = box(1, 10, 1.0, 0.5)
    assert lib.or32_aabb_intersects(s1, s2)
    # the 0.01 padding: a gap of 0.0099 still overlaps, 0.0101 does not
    assert lib.or32_aabb_intersects(box(0, 0, 0.5, 0.5), box(1.0099, 0, 0.5, 0.5))
    assert not lib.or32_aabb_intersects(box(0, 0, 0.5, 0.5), box(1.0101, 0, 0.5, 0.5))


def test_spatial_hash_function(oracle):
    lib = oracle.load()
    lib.or_hash.restype = ctypes.c_uint64
    lib.or_hash.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64]
    M = 1 << 64
    for (t, x, y) in [(10, 0, 0), (422, 3, -7), (8194, -123456, 98765), (2, -1, -1), (1000003, 2**31, -2**31)]:
        h = (((x * 92837111) % M) ^ ((y * 689287499) % M)) % t
        assert lib.or_hash(t, x, y) == h


def test_oracle_pool_equals_in_process():
    """The spawned-process oracle runner used by the full-batch GPU check
    gives the in-process oracle's bits."""
    from helpers import bits_equal
    from oracle_pool import oracle_bodies, run_worlds
    from resolve2d_amd import scenes
    specs = [scenes.batched_world(1000 + i, n_min=50, n_max=120) for i in range(6)]
    ref = oracle_bodies(specs, 10, scenes.DT, 4, 4, 3)
    loc = run_worlds((specs, 10, scenes.DT, 4, 4))
    assert len(ref) == len(specs) and all(bits_equal(a, b) for a, b in zip(ref, loc))
