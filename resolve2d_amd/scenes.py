"""Seeded scene generators (synthetic inputs for tests and the benchmark).

A scene is a list of operations; ids are assigned in creation order from 0,
exactly as EntityFactory.appendBody does (src/core/lib.zig:66-71), so the
generators can refer to the bodies they create.  Every number is rounded to
f32 here so that every engine receives identical inputs.

Operations:
  ("disc", px, py, vx, vy, angle, omega, mu, mass_kind, mass_value, radius)
  ("rect", px, py, vx, vy, angle, omega, mu, mass_kind, mass_value, width, height)
  ("static", id, flag)          ("remove", id)        ("gravity", g)
  ("spring", id, k, c, anchor_x, anchor_y)            ("exclude", id1, id2)
  ("joint", type, pmax, pmin, beta, id1, id2, r1x, r1y, r2x, r2y, target, tx, ty)
  ("hash", cell_width, table_mult)   -- tuned spatial-hash mode
mass_kind 0 = density, 1 = mass.  Joint types: 0 distance, 1 offset
distance, 2 fixed position, 3 motor.
"""
import math

import numpy as np

INF = float("inf")
DT = 1.0 / 60.0
SUB_STEPS = 4
ITERS = 4


def f32(x) -> float:
    return float(np.float32(x))


class _Builder:
    def __init__(self):
        self.ops = []
        self.n = 0

    def disc(self, x, y, r, mu=0.5, density=1.0, mass=None, vx=0.0, vy=0.0, angle=0.0, omega=0.0):
        kind, val = (1, mass) if mass is not None else (0, density)
        self.ops.append(("disc", f32(x), f32(y), f32(vx), f32(vy), f32(angle), f32(omega), f32(mu), kind, f32(val),
                         f32(r)))
        self.n += 1
        return self.n - 1

    def rect(self, x, y, w, h, mu=0.5, density=1.0, mass=None, vx=0.0, vy=0.0, angle=0.0, omega=0.0):
        kind, val = (1, mass) if mass is not None else (0, density)
        self.ops.append(("rect", f32(x), f32(y), f32(vx), f32(vy), f32(angle), f32(omega), f32(mu), kind, f32(val),
                         f32(w), f32(h)))
        self.n += 1
        return self.n - 1

    def static(self, i, flag=True):
        self.ops.append(("static", i, bool(flag)))

    def gravity(self, g):
        self.ops.append(("gravity", f32(g)))

    def spring(self, i, k, c, ax, ay):
        self.ops.append(("spring", i, f32(k), f32(c), f32(ax), f32(ay)))

    def joint(self, type_, id1, id2=0, r1=(0.0, 0.0), r2=(0.0, 0.0), target=0.0, tpos=(0.0, 0.0),
              pmax=INF, pmin=-INF, beta=10.0):
        self.ops.append(("joint", type_, f32(pmax), f32(pmin), f32(beta), id1, id2, f32(r1[0]), f32(r1[1]),
                         f32(r2[0]), f32(r2[1]), f32(target), f32(tpos[0]), f32(tpos[1])))

    def exclude(self, a, b):
        self.ops.append(("exclude", a, b))


def pyramid(rows: int = 20, mu: float = 0.5, g: float = 9.81):
    """Box pyramid: rows*(rows+1)/2 unit boxes (density 1) touching exactly,
    on a static 100 x 1 ground rectangle at (0, -0.5)."""
    b = _Builder()
    b.gravity(g)
    ground = b.rect(0.0, -0.5, 100.0, 1.0, mu=mu)
    b.static(ground)
    for r in range(rows):
        for i in range(rows - r):
            b.rect(i - (rows - 1 - r) / 2.0, 0.5 + r, 1.0, 1.0, mu=mu)
    return b.ops


def max_extent(ops) -> float:
    """Largest side of a dynamic body's axis-aligned box as created: 2 r for a
    disc, |w cos a| + |h sin a| or |w sin a| + |h cos a| for a rectangle."""
    static = {op[1] for op in ops if op[0] == "static" and op[2]}
    best, i = 0.0, 0
    for op in ops:
        if op[0] not in ("disc", "rect"):
            continue
        if i not in static:
            if op[0] == "disc":
                best = max(best, 2.0 * op[10])
            else:
                c, s = abs(math.cos(op[5])), abs(math.sin(op[5]))
                w, h = op[10], op[11]
                best = max(best, w * c + h * s, w * s + h * c)
        i += 1
    return best


def tuned_cell(ops) -> float:
    """Cell width of the tuned spatial hash: twice the largest body extent."""
    return f32(2.0 * max_extent(ops))


def mixed(n: int = 4096, seed: int = 1, width: float = 100.0, y0: float = 1.0, height: float = 200.0,
          ground_w: float = 120.0, mu: float = 0.4, g: float = 9.81, hash_cell=None):
    """n bodies alternating disc / rectangle: disc radius U[0.2, 0.6]; rectangle
    sides U[0.3, 1.0], angle U[0, 2pi); positions uniform in a width x height box
    above a static ground; density 1.  hash_cell selects the tuned hash mode
    with that cell width (2N buckets); "auto" takes twice the largest body
    extent (tuned_cell)."""
    if hash_cell == "auto":
        base = mixed(n, seed, width, y0, height, ground_w, mu, g)
        return [("hash", tuned_cell(base), 2)] + base
    rng = np.random.default_rng(seed)
    b = _Builder()
    if hash_cell is not None:
        b.ops.append(("hash", f32(hash_cell), 2))
    b.gravity(g)
    ground = b.rect(0.0, -0.5, ground_w, 1.0, mu=mu)
    b.static(ground)
    for i in range(n):
        x = rng.uniform(-width / 2, width / 2)
        y = rng.uniform(y0, y0 + height)
        if i % 2 == 0:
            b.disc(x, y, rng.uniform(0.2, 0.6), mu=mu)
        else:
            b.rect(x, y, rng.uniform(0.3, 1.0), rng.uniform(0.3, 1.0), mu=mu, angle=rng.uniform(0.0, 2 * math.pi))
    return b.ops


def batched_world(seed: int, n_min: int = 50, n_max: int = 500, joints_max: int = 20, springs_max: int = 10):
    """One world of the batched workload: N ~ U{n_min..n_max} mixed bodies (as
    `mixed`, in a box whose area grows with N), 0..joints_max distance joints
    between random body pairs (rest length = initial distance, default
    parameters) and 0..springs_max anchored springs (k ~ U[10, 100], c = 0.1,
    anchored at the body's start position)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(n_min, n_max + 1))
    width = 40.0
    height = max(10.0, n * 0.12)
    b = _Builder()
    b.gravity(9.81)
    ground = b.rect(0.0, -0.5, width + 20.0, 1.0, mu=0.4)
    b.static(ground)
    pos = []
    for i in range(n):
        x = f32(rng.uniform(-width / 2, width / 2))
        y = f32(rng.uniform(1.0, 1.0 + height))
        pos.append((x, y))
        if i % 2 == 0:
            b.disc(x, y, rng.uniform(0.2, 0.6), mu=0.4)
        else:
            b.rect(x, y, rng.uniform(0.3, 1.0), rng.uniform(0.3, 1.0), mu=0.4, angle=rng.uniform(0.0, 2 * math.pi))
    nj = int(rng.integers(0, joints_max + 1))
    for _ in range(nj):
        a, c = rng.choice(n, size=2, replace=False)
        (xa, ya), (xc, yc) = pos[a], pos[c]
        d = f32(math.sqrt((xc - xa) ** 2 + (yc - ya) ** 2))
        b.joint(0, int(a) + 1, int(c) + 1, target=d)
    ns = int(rng.integers(0, springs_max + 1))
    for _ in range(ns):
        a = int(rng.integers(0, n))
        b.spring(a + 1, rng.uniform(10.0, 100.0), 0.1, pos[a][0], pos[a][1])
    return b.ops


def many_boxes_demo():
    """The reference's scene 0_3 (src/examples/0_3_many_boxes.zig:12-61)."""
    b = _Builder()
    b.gravity(9.82)
    g = b.rect(20.0, -5.0, 1000.0, 10.0, mu=0.3, density=5.0)
    b.static(g)
    b.rect(-40.0, 10.0, 8.0, 8.0, mu=0.3, mass=100.0, vx=70.0, vy=10.0, omega=-6.0)
    b.rect(20.0, 5.0, 4.0, 1.0, mu=0.3, density=5.0)
    for x in range(10, 30):
        xf = 2.0 * x
        for y in range(10, 36):
            if y % 2 == 0:
                b.rect(xf, float(y), 1.0, 1.0, mu=0.3, density=5.0)
            else:
                b.disc(xf, float(y), 0.5, mu=0.3, density=5.0)
    return b.ops


def car_demo():
    """The car of src/examples/utils.zig:6-54 on the first platform of
    src/examples/0_1_car_platformer.zig:12-36: all four joint types and
    excluded pairs."""
    b = _Builder()
    b.gravity(9.82)
    p1 = b.rect(0.0, -100.0, 1000.0, 20.0, mu=0.3)
    b.static(p1)
    p2 = b.rect(0.0, 0.0, 40.0, 10.0, mu=0.3)
    b.static(p2)
    p3 = b.rect(-19.0, 15.0, 2.0, 20.0, mu=0.3)
    b.static(p3)
    # utils.car(fac, (5, 10))
    px, py = 5.0, 10.0
    tx, ty = px - 5.0, py - 10.0
    bh = b.rect(px, py, 5.0, 0.9, mu=0.4)
    rad = 1.0
    wl = b.disc(tx + 3.5, ty + 9.8 - rad, rad, mu=1.5)
    wr = b.disc(tx + 6.5, ty + 9.8 - rad, rad, mu=1.5)
    dist_car = f32(math.hypot(f32(tx + 3.5) - px, f32(ty + 9.8 - rad) - py))
    bh2 = b.rect(tx + 5.25, ty + 10.8, 1.2, 0.5, mu=0.5)
    P = dict(beta=14.0, pmin=-2.0, pmax=2.0)
    b.joint(1, wl, bh, r1=(0.0, 0.0), r2=(-1.5, 0.0), target=rad + 0.2, **P)
    b.joint(1, wr, bh, r1=(0.0, 0.0), r2=(1.5, 0.0), target=rad + 0.2, **P)
    b.joint(0, wl, wr, target=3.0, **P)
    b.joint(0, wl, bh, target=dist_car, **P)
    b.joint(0, wr, bh, target=dist_car, **P)
    b.exclude(bh, wl)
    b.exclude(bh, wr)
    b.exclude(bh, bh2)
    b.joint(1, bh, bh2, r1=(0.25, 0.0), r2=(-1.0, 1.0), target=f32(math.hypot(px + 0.25 - (tx + 5.25 - 1.0),
                                                                             py - (ty + 10.8 + 1.0))))
    b.joint(1, bh, bh2, r1=(0.25, 0.0), r2=(1.0, 1.0), target=f32(math.hypot(px + 0.25 - (tx + 5.25 + 1.0),
                                                                            py - (ty + 10.8 + 1.0))))
    b.joint(0, bh, bh2, beta=100.0, target=f32(math.hypot(px - (tx + 5.25), py - (ty + 10.8))))
    # a motor on the right wheel and a fixed-position anchor keep every joint type in use
    b.joint(3, wr, target=-4.0, pmax=50.0, pmin=-50.0)
    anchor = b.rect(-15.0, 9.0, 0.6, 0.4, mu=0.5)
    b.joint(2, anchor, tpos=(-15.0, 9.5), beta=5.0)
    return b.ops


# ------------------------------------------------------------- served scenes
def served_arrays(seed: int, n: int):
    """A served scene of n mixed bodies with the geometry and materials of
    `mixed` (discs r ~ U[0.2, 0.6], rectangles w, h ~ U[0.3, 1.0], angle ~
    U[0, 2pi), density 1, friction 0.4, alternating disc / rectangle), drawn
    vectorised so that scenes of thousands of bodies are cheap to generate.
    The box scales with n at the density of the 4,096-body scene (100 x 200 m):
    width = max(8, 100 sqrt(n / 4096)), height = 2 width.  Returns float32 /
    int32 arrays in creation order (the ground is created first, as id 0)."""
    rng = np.random.default_rng(seed)
    width = max(8.0, 100.0 * math.sqrt(n / 4096.0))
    height = 2.0 * width
    x = rng.uniform(-width / 2, width / 2, n).astype(np.float32)
    y = rng.uniform(1.0, 1.0 + height, n).astype(np.float32)
    r = rng.uniform(0.2, 0.6, n).astype(np.float32)
    w = rng.uniform(0.3, 1.0, n).astype(np.float32)
    h = rng.uniform(0.3, 1.0, n).astype(np.float32)
    ang = rng.uniform(0.0, 2 * math.pi, n).astype(np.float32)
    kind = (np.arange(n) % 2).astype(np.int32)            # 0 disc, 1 rectangle
    rect = kind == 1
    return dict(n=n, width=np.float32(width), kind=kind, px=x, py=y,
                angle=np.where(rect, ang, np.float32(0.0)).astype(np.float32),
                dim_a=np.where(rect, w, r).astype(np.float32), dim_b=np.where(rect, h, np.float32(0.0)).astype(np.float32))


def served_ops(seed: int, n: int):
    """The scene of `served_arrays` as a list of scene operations (for the
    oracle and small tests)."""
    a = served_arrays(seed, n)
    b = _Builder()
    b.gravity(9.81)
    ground = b.rect(0.0, -0.5, float(a["width"]) + 20.0, 1.0, mu=0.4)
    b.static(ground)
    for i in range(n):
        if a["kind"][i] == 0:
            b.disc(a["px"][i], a["py"][i], a["dim_a"][i], mu=0.4)
        else:
            b.rect(a["px"][i], a["py"][i], a["dim_a"][i], a["dim_b"][i], mu=0.4, angle=a["angle"][i])
    return b.ops


def build_served(world, seed: int, n: int):
    """Build the scene of `served_arrays` into an engine world with one bulk
    call (GpuWorld.make_bodies); the result equals applying `served_ops`."""
    a = served_arrays(seed, n)
    world.make_gravity(f32(9.81))
    g = world.make_rect(0.0, -0.5, 0.0, 0.0, 0.0, 0.0, f32(0.4), 0, 1.0, f32(float(a["width"]) + 20.0), 1.0)
    world.set_static(g, True)
    z = np.zeros(n, dtype=np.float32)
    world.make_bodies(a["kind"], a["px"], a["py"], z, z, a["angle"], z, np.full(n, f32(0.4), dtype=np.float32),
                      np.zeros(n, dtype=np.int32), np.ones(n, dtype=np.float32), a["dim_a"], a["dim_b"])
    return world


def poisson_stream(seed: int, rate: float, count: int, n_min: int = 16, n_max: int = 8192):
    """Arrival times (s, from 0) of a Poisson stream of `count` scenes at
    `rate` scenes/s, with sizes log-uniform in [n_min, n_max] and one scene
    seed each: a list of (t, n, scene_seed)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    n = np.exp(rng.uniform(math.log(n_min), math.log(n_max), count)).round().astype(int)
    seeds = rng.integers(0, 2 ** 31, count)
    return [(float(t[i]), int(n[i]), int(seeds[i])) for i in range(count)]


def large_constrained(seed: int = 4, n_bodies: int = 65536, ropes: int = 100, links: int = 20, motors: int = 64,
                      size: float = 600.0):
    """The large constrained world: a size x size m domain walled by four
    static boxes (floor, ceiling, left, right; 10 m thick), gravity 9.81, and
    n_bodies dynamic bodies in total:
      - `ropes` ropes of `links` distance joints each: links + 1 discs of
        radius 0.2 spaced 0.8 m apart horizontally, rest length = initial
        distance (default parameters);
      - `motors` wheels (discs of radius 1, density 1) driven by motor
        joints, target angular velocity ~ U[-5, 5] rad/s, power cap +-50
        (the reference's motor adds beta (w* - w) dt to the angular momentum,
        which overshoots unless beta dt / I < 1: wheels of inertia pi / 2
        keep it at 0.03);
      - the remaining bodies mixed as in `mixed` (discs r ~ U[0.2, 0.6],
        rectangles w, h ~ U[0.3, 1.0], angle ~ U[0, 2pi), density 1,
        friction 0.4) at uniform positions inside the walls.
    Body ids exceed the reference's u16 range (65,540 bodies)."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    b.gravity(9.81)
    h = size / 2.0
    t = 10.0
    walls = [(0.0, -t / 2, size + 2 * t, t), (0.0, size + t / 2, size + 2 * t, t),
             (-h - t / 2, size / 2, t, size), (h + t / 2, size / 2, t, size)]
    for (x, y, w, hh) in walls:
        b.static(b.rect(x, y, w, hh, mu=0.4))
    inner = h - 5.0
    for r in range(ropes):
        x0 = rng.uniform(-inner, inner - 0.8 * links)
        y0 = rng.uniform(5.0, size - 5.0)
        ids = [b.disc(x0 + 0.8 * k, y0, 0.2, mu=0.4) for k in range(links + 1)]
        for k in range(links):
            b.joint(0, ids[k], ids[k + 1], target=f32(0.8 * (k + 1) + x0) - f32(0.8 * k + x0))
    omega = rng.uniform(-5.0, 5.0, motors)
    for k in range(motors):
        wheel = b.disc(rng.uniform(-inner, inner), rng.uniform(5.0, size - 5.0), 1.0, mu=0.4)
        b.joint(3, wheel, target=omega[k], pmax=50.0, pmin=-50.0)
    rest = n_bodies - ropes * (links + 1) - motors
    x = rng.uniform(-inner, inner, rest)
    y = rng.uniform(5.0, size - 5.0, rest)
    rad = rng.uniform(0.2, 0.6, rest)
    w = rng.uniform(0.3, 1.0, rest)
    hgt = rng.uniform(0.3, 1.0, rest)
    ang = rng.uniform(0.0, 2 * math.pi, rest)
    for i in range(rest):
        if i % 2 == 0:
            b.disc(x[i], y[i], rad[i], mu=0.4)
        else:
            b.rect(x[i], y[i], w[i], hgt[i], mu=0.4, angle=ang[i])
    return b.ops
