"""The Zig-shaped host API (resolve2d_amd.solver: Solver, EntityFactory,
BodyHandle, RigidBody.props; src/root.zig:4-10, src/core/lib.zig:13-315)
driven as the reference's callers drive it, bit for bit against the oracle.

Scenes are applied through EntityFactory.make* / bodyHandle(id).body_unwrap()
/ removeRigidBody, stepped with Solver.process, edited between steps through
props (as demos/native/src/main.zig:114-135 does), cleared with Solver.clear
(lib.zig:181-187) and rebuilt."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def apply_via_factory(ops, solver):
    """Scene operations (resolve2d_amd.scenes) through the Zig-named API."""
    from resolve2d_amd.solver import BodyOptions, DiscOptions, Parameters, RectangleOptions
    fac = solver.entityFactory()
    for op in ops:
        kind = op[0]
        if kind in ("disc", "rect"):
            px, py, vx, vy, ang, om, mu, mk, mv = op[1:10]
            bo = BodyOptions(pos=(px, py), vel=(vx, vy), angle=ang, omega=om, mu=mu,
                             density=mv if mk == 0 else None, mass=mv if mk == 1 else None)
            if kind == "disc":
                fac.makeDiscBody(bo, DiscOptions(radius=op[10]))
            else:
                fac.makeRectangleBody(bo, RectangleOptions(width=op[10], height=op[11]))
        elif kind == "static":
            solver.bodyHandle(op[1]).body_unwrap().static = op[2]
        elif kind == "remove":
            solver.removeRigidBody(op[1])
        elif kind == "gravity":
            fac.makeDownwardsGravity(op[1])
        elif kind == "spring":
            fac.makeSpring(solver.bodyHandle(op[1]), op[2], op[3], (op[4], op[5]))
        elif kind == "joint":
            t, pmax, pmin, beta, i1, i2, r1x, r1y, r2x, r2y, target, tx, ty = op[1:]
            p = Parameters(power_max=pmax, power_min=pmin, beta=beta)
            h1, h2 = solver.bodyHandle(i1), solver.bodyHandle(i2)
            if t == 0:
                fac.makeDistanceJoint(p, h1, h2, target)
            elif t == 1:
                fac.makeOffsetDistanceJoint(p, h1, h2, (r1x, r1y), (r2x, r2y), target)
            elif t == 2:
                fac.makeFixedPositionJoint(p, h1, (tx, ty))
            else:
                fac.makeMotorJoint(p, h1, target)
        elif kind == "exclude":
            fac.excludeCollisionPair(solver.bodyHandle(op[1]), solver.bodyHandle(op[2]))
        else:
            raise ValueError(kind)


def solver_bodies(solver):
    """Solver states as the oracle's body rows (id, type, static, pose, ...)."""
    from resolve2d_amd.engine import BODY_COLS
    st = solver.states()
    out = np.zeros((len(st), BODY_COLS))
    for i, b in enumerate(st):
        out[i] = [b.id, b.type, b.is_static, b.pos_x, b.pos_y, b.mom_x, b.mom_y, b.force_x, b.force_y, b.mass,
                  b.angle, b.ang_momentum, b.torque, b.inertia, b.mu, b.aabb_x, b.aabb_y, b.aabb_half_width,
                  b.aabb_half_height]
    return out


def _compare(solver, o, where):
    a, b = solver_bodies(solver), o.bodies()
    assert bits_equal(a, b), (where, first_diff(a, b))


@pytest.mark.parametrize("name", ["pyramid", "car"])
def test_factory_built_scene_bit_exact(gpu_ctx, oracle, name):
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import apply_scene
    from resolve2d_amd.solver import Solver
    ops = scenes.pyramid(rows=12) if name == "pyramid" else scenes.car_demo()
    s = Solver.init(4.0, 2, ctx=gpu_ctx)
    o = oracle.OracleWorld("f32")
    apply_via_factory(ops, s)
    apply_scene(ops, o)
    assert s.num_bodies() == len(o.bodies())
    assert [s.body_id_at(i) for i in range(s.num_bodies())] == [int(x) for x in o.bodies()[:, 0]]
    _compare(s, o, "built")
    for step in range(1, 61):
        s.process(DT, 4, 4)
        assert o.process(DT, 4, 4) == 0
        if step % 20 == 0:
            _compare(s, o, step)
    s.deinit()


def test_props_edits_removal_and_handles(gpu_ctx, oracle):
    """Between steps: write momentum / angular momentum / torque through
    props, toggle static through body_unwrap(), remove a body (swapRemove:
    the last body takes its iteration slot); a removed id's handle gives
    None, and removing it again raises InvalidRigidBodyId."""
    from resolve2d_amd import scenes
    from resolve2d_amd._native import InvalidRigidBodyId
    from resolve2d_amd.engine import apply_scene
    from resolve2d_amd.solver import Solver
    # the car scene plus a few free boxes (pyramid rows without their ground)
    ops = scenes.car_demo() + [op for op in scenes.pyramid(rows=4) if op[0] == "rect"][1:]
    s = Solver.init(ctx=gpu_ctx)
    o = oracle.OracleWorld("f32")
    apply_via_factory(ops, s)
    apply_scene(ops, o)
    for _ in range(10):
        s.process(DT, 4, 4)
        o.process(DT, 4, 4)
    n = s.num_bodies()
    # a body no joint names (removing a jointed one makes the next step fail
    # with InvalidRigidBodyId, as in the reference)
    jointed = {op[5] for op in ops if op[0] == "joint"} | {op[6] for op in ops if op[0] == "joint"}
    slot = next(i for i in range(3, n - 2) if s.body_id_at(i) not in jointed and not
                s.bodyHandle(s.body_id_at(i)).body_unwrap().static)
    victim = s.body_id_at(slot)
    last = s.body_id_at(n - 1)
    edit = s.body_id_at(n - 2)
    b = s.bodyHandle(edit).body_unwrap()
    b.props.momentum = (1.5, -0.25)
    b.props.ang_momentum = 0.75
    b.props.torque = 2.0
    row = [r for r in o.bodies() if int(r[0]) == edit][0]
    o.set_body(edit, row[3], row[4], np.float32(1.5), np.float32(-0.25), row[7], row[8], row[10],
               np.float32(0.75), np.float32(2.0))
    pinned = s.body_id_at(slot + 1)
    s.bodyHandle(pinned).body_unwrap().static = True
    o.set_static(pinned, True)
    assert s.bodyHandle(pinned).body_unwrap().static
    s.removeRigidBody(victim)
    o.remove_body(victim)
    assert s.body_id_at(slot) == last
    assert s.bodyHandle(victim).body() is None
    with pytest.raises(InvalidRigidBodyId):
        s.removeRigidBody(victim)
    _compare(s, o, "edited")
    for _ in range(30):
        s.process(DT, 4, 4)
        assert o.process(DT, 4, 4) == 0
    _compare(s, o, "after edits")
    s.deinit()


def test_clear_then_rebuild_bit_exact(gpu_ctx, oracle):
    """Solver.clear (lib.zig:181-187) drops bodies, joints, force generators,
    manifolds and excluded pairs; ids keep counting.  A scene built after the
    clear steps exactly as in the oracle after its clear."""
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import apply_scene
    from resolve2d_amd.solver import Solver
    s = Solver.init(ctx=gpu_ctx)
    o = oracle.OracleWorld("f32")
    first = scenes.car_demo()
    apply_via_factory(first, s)
    apply_scene(first, o)
    for _ in range(5):
        s.process(DT, 4, 4)
        o.process(DT, 4, 4)
    s.clear()
    o.clear()
    assert s.num_bodies() == 0 and len(s.manifolds()) == 0
    s.process(DT, 4, 4)            # an empty world steps
    assert o.process(DT, 4, 4) == 0
    # the second scene's ids continue after the first's
    second = scenes.pyramid(rows=8)
    offset = int(o.lib.or32_current_id(o.h))
    shifted = []
    for op in second:
        if op[0] == "static":
            shifted.append(("static", op[1] + offset, op[2]))
        else:
            shifted.append(op)
    apply_via_factory(shifted, s)
    apply_scene(shifted, o)
    assert s.body_id_at(0) == offset
    for step in range(40):
        s.process(DT, 4, 4)
        assert o.process(DT, 4, 4) == 0
    _compare(s, o, "rebuilt")
    mg = np.array([[m.ref_id, m.inc_id] for m in s.manifolds()])
    mo = o.manifolds()
    assert len(mg) == len(mo) and (mg == mo[:, :2]).all()
    s.deinit()
