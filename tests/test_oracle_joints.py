"""Pins of the oracle's joints (dynamics/joints/b2*joint.d) by closed forms and
invariants the oracle does not compute itself:

  distance   a rigid pendulum's period 2 pi sqrt(L / g) and its constant length; a
             soft one's spring period
  revolute   a hinged bar's physical-pendulum period 2 pi sqrt(I_pivot / (m g d)) and
             its fixed anchor; the motor's speed and torque limit; the angle limit
  prismatic  free fall along an inclined axis at g sin(theta) with no motion across
             it; the lower translation limit; the motor's speed
  weld       a welded pair keeps its relative pose through a fall and a landing
  rope       a slack rope exerts nothing; a taut one holds the length
  wheel      the suspension spring's period 1 / f
  friction   a sliding body decelerates at exactly maxForce / m
  motor      the body is driven to the joint's linear and angular offset
  mouse      the body is pulled to the target
  pulley     lengthA + ratio lengthB stays constant; the heavier side goes down
  filter     collideConnected decides whether jointed bodies get a contact
  islands    a joint joins its bodies in one island: they fall asleep together and
             wake each other

A plausible restatement mistake (a sign, a swapped body, a missing term) moves one
of these by far more than its tolerance.
"""
import math

import numpy as np
import pytest

from dbox_amd import scenes
from oracle import pyoracle as po

DT = 1.0 / 60.0


def world(bodies, fixtures, joints, gravity=(0.0, -10.0)):
    b = scenes.body_defs(len(bodies))
    for i, (t, x, y, a) in enumerate(bodies):
        b[i]["type"], b[i]["px"], b[i]["py"], b[i]["angle"] = t, x, y, a
    f = np.array(fixtures, scenes.FIXTURE_DTYPE) if fixtures else scenes.fixture_defs(0)
    return scenes.Scene(b, f, gravity=gravity, joints=np.array(joints, scenes.JOINT_DTYPE))


def circle(body, r, density=1.0):
    f = scenes.fixture_defs(1)[0]
    f["body"], f["shape"], f["radius"], f["density"] = body, scenes.SHAPE_CIRCLE, r, density
    return f


def box(body, hx, hy, density=1.0, friction=0.6):
    f = scenes.fixture_defs(1)[0]
    f["body"], f["shape"], f["density"], f["friction"] = body, scenes.SHAPE_BOX, density, friction
    f["p"][:2] = (hx, hy)
    return f


def joint(kind, a, b, **kw):
    j = scenes.joint_defs(1, kind)[0]
    j["body_a"], j["body_b"] = a, b
    for k, v in kw.items():
        j[k] = v
    return j


def half_periods(xs):
    """Times (in steps) of the sign changes of xs, interpolated linearly."""
    xs = np.asarray(xs, np.float64)
    i = np.nonzero(np.sign(xs[:-1]) != np.sign(xs[1:]))[0]
    return i + xs[i] / (xs[i] - xs[i + 1])


def run(scene, steps, **kw):
    o = po.OracleWorld(scene, allow_sleep=False, **kw)
    out = []
    for _ in range(steps):
        o.step(DT)
        out.append(o.bodies().copy())
    return o, out


# ---------------------------------------------------------------- distance
def test_distance_pendulum_period_and_length():
    L, th = 2.0, 0.08
    sc = world([(0, 0, 10, 0), (2, L * math.sin(th), 10 - L * math.cos(th), 0)], [circle(1, 0.05)],
               [joint(scenes.DISTANCE, 0, 1, length=L)])
    o, st = run(sc, 700)
    x = [s["cx"][1] for s in st]
    t = half_periods(x)
    period = 2 * np.mean(np.diff(t)) * DT
    expect = 2 * math.pi * math.sqrt(L / 10.0) * (1 + th * th / 16)
    assert period == pytest.approx(expect, rel=2e-3)
    d = [math.hypot(s["cx"][1], s["cy"][1] - 10) for s in st]
    assert max(abs(v - L) for v in d) < 1e-4


def test_soft_distance_joint_spring_period():
    # zero gravity, a 1 kg bob on a 2 Hz distance spring stretched 0.3 m: it oscillates
    # about the rest length at about 2 Hz (the Box2D soft constraint is implicit, so the
    # frequency is slightly lower than the nominal one at 60 Hz)
    r = 1.0 / math.sqrt(math.pi)  # unit mass at density 1
    sc = world([(0, 0, 0, 0), (2, 1.3, 0, 0)], [circle(1, r)],
               [joint(scenes.DISTANCE, 0, 1, length=1.0, frequency_hz=2.0, damping_ratio=0.0)], gravity=(0, 0))
    o, st = run(sc, 300)
    x = np.array([s["cx"][1] for s in st]) - 1.0
    period = 2 * np.mean(np.diff(half_periods(x))) * DT
    assert 0.5 < period < 0.56
    assert np.abs(x).max() < 0.31  # no energy gain


# ---------------------------------------------------------------- revolute
def test_revolute_physical_pendulum_period_and_anchor():
    # a 2 m x 0.2 m bar hinged at one end, 5 degrees from hanging straight down
    th = math.radians(5)
    d = 1.0
    cx, cy = d * math.sin(th), 10 - d * math.cos(th)
    sc = world([(0, 0, 10, 0), (2, cx, cy, th - math.pi / 2)], [box(1, 1.0, 0.1)],
               [joint(scenes.REVOLUTE, 0, 1, local_anchor_b=(-1.0, 0.0))])
    o, st = run(sc, 700)
    x = [s["cx"][1] for s in st]
    period = 2 * np.mean(np.diff(half_periods(x))) * DT
    i_pivot = (4.0 + 0.04) / 12.0 + d * d  # per unit mass
    expect = 2 * math.pi * math.sqrt(i_pivot / (10.0 * d)) * (1 + th * th / 16)
    assert period == pytest.approx(expect, rel=2e-3)
    for s in st:  # the hinge point of the bar stays at the static anchor
        px = s["px"][1] + s["qc"][1] * -1.0
        py = s["py"][1] + s["qs"][1] * -1.0
        assert math.hypot(px, py - 10) < 2e-4


def test_revolute_motor_speed_and_torque_limit():
    r = 0.5
    wheel = circle(1, r)
    inertia = (math.pi * r * r) * r * r / 2
    # enough torque: the wheel turns at the motor speed after one step
    sc = world([(0, 0, 0, 0), (2, 0, 0, 0)], [wheel],
               [joint(scenes.REVOLUTE, 0, 1, enable_motor=1, motor_speed=3.0, max_motor=1e3)], gravity=(0, 0))
    o, st = run(sc, 3)
    assert st[0]["w"][1] == pytest.approx(3.0, rel=1e-5)
    # limited torque tau: w grows by tau dt / I per step
    tau = 0.05
    sc = world([(0, 0, 0, 0), (2, 0, 0, 0)], [wheel],
               [joint(scenes.REVOLUTE, 0, 1, enable_motor=1, motor_speed=100.0, max_motor=tau)], gravity=(0, 0))
    o, st = run(sc, 30)
    for n in (1, 10, 30):
        assert st[n - 1]["w"][1] == pytest.approx(n * tau * DT / inertia, rel=2e-4)


def test_revolute_limit_stops_the_bar():
    # the bar starts horizontal and swings down; the limit stops it at -0.5 rad from
    # its reference angle
    sc = world([(0, 0, 10, 0), (2, 1.0, 10, 0)], [box(1, 1.0, 0.1)],
               [joint(scenes.REVOLUTE, 0, 1, local_anchor_b=(-1.0, 0.0), enable_limit=1, lower=-0.5, upper=0.5)])
    o, st = run(sc, 240)
    a = np.array([s["a"][1] for s in st])
    assert a.min() > -0.5 - 0.05  # overshoot bounded by the Baumgarte-free velocity limit
    assert abs(a[-1] - (-0.5)) < 2 * math.radians(2)  # resting at the limit within 2 angular slops
    assert abs(st[-1]["w"][1]) < 1e-3


# ---------------------------------------------------------------- prismatic
def test_prismatic_slides_along_its_axis():
    th = math.radians(30)
    axis = (math.cos(th), -math.sin(th))  # down a 30 degree slope
    sc = world([(0, 0, 0, 0), (2, 0, 0, 0)], [box(1, 0.5, 0.5)],
               [joint(scenes.PRISMATIC, 0, 1, local_axis_a=axis)])
    o, st = run(sc, 60)
    n = 60
    # semi-implicit Euler along the axis: s_n = h^2 a n (n + 1) / 2 with a = g sin(theta)
    s_expect = DT * DT * 10.0 * math.sin(th) * n * (n + 1) / 2
    p = np.array([st[-1]["cx"][1], st[-1]["cy"][1]])
    s_along = p @ np.array(axis)
    s_across = p @ np.array([-axis[1], axis[0]])
    assert s_along == pytest.approx(s_expect, rel=1e-4)
    assert abs(s_across) < 1e-5 and abs(st[-1]["a"][1]) < 1e-5


def test_prismatic_limit_and_motor():
    sc = world([(0, 0, 0, 0), (2, 0, 0, 0)], [box(1, 0.5, 0.5)],
               [joint(scenes.PRISMATIC, 0, 1, local_axis_a=(0.0, 1.0), enable_limit=1, lower=-1.0, upper=1.0)])
    o, st = run(sc, 120)
    y = np.array([s["cy"][1] for s in st])
    assert y.min() > -1.0 - 0.03 and abs(y[-1] + 1.0) < 0.01
    sc = world([(0, 0, 0, 0), (2, 0, 0, 0)], [box(1, 0.5, 0.5)],
               [joint(scenes.PRISMATIC, 0, 1, local_axis_a=(1.0, 0.0), enable_motor=1, motor_speed=2.0,
                      max_motor=1e4)], gravity=(0, 0))
    o, st = run(sc, 2)
    assert st[0]["vx"][1] == pytest.approx(2.0, rel=1e-5)
    assert abs(st[0]["vy"][1]) < 1e-6


# ---------------------------------------------------------------- weld, rope
def test_weld_keeps_the_relative_pose():
    ground = scenes.fixture_defs(1)[0]
    scenes._edge(ground, 0, -20, 0, 20, 0, 0.6)
    sc = world([(0, 0, 0, 0), (2, 0, 3, 0.3), (2, 1.2, 3.4, 0.3)], [ground, box(1, 0.5, 0.2), box(2, 0.5, 0.2)],
               [joint(scenes.WELD, 1, 2, local_anchor_a=(0.6, 0.2), local_anchor_b=(-0.6, -0.2))])
    o, st = run(sc, 180)
    rel0 = None
    for s in st:
        da = s["a"][2] - s["a"][1]
        c, sn = math.cos(s["a"][1]), math.sin(s["a"][1])
        dx, dy = s["cx"][2] - s["cx"][1], s["cy"][2] - s["cy"][1]
        rel = (c * dx + sn * dy, -sn * dx + c * dy, da)
        if rel0 is None:
            rel0 = rel
        assert max(abs(rel[k] - rel0[k]) for k in range(3)) < 2e-3
    assert st[-1]["cy"][1] < 1.0  # it did fall and land


def test_rope_slack_and_taut():
    # slack: the bob falls freely until the rope is straight
    sc = world([(0, 0, 10, 0), (2, 0, 9, 0)], [circle(1, 0.1)], [joint(scenes.ROPE, 0, 1, length=3.0)])
    o, st = run(sc, 60)
    for n in range(1, 12):  # 11 steps fall 0.2 m... still slack
        assert st[n - 1]["cy"][1] == pytest.approx(9.0 - DT * DT * 10.0 * n * (n + 1) / 2, abs=1e-5)
    d = [10 - s["cy"][1] for s in st]
    assert max(d) < 3.0 + 0.005 + 1e-4 and d[-1] > 2.99


# ---------------------------------------------------------------- wheel, friction, motor, mouse, pulley
def test_wheel_suspension_period():
    r = 1.0 / math.sqrt(math.pi)
    sc = world([(0, 0, 0, 0), (2, 0, 0.3, 0)], [circle(1, r)],
               [joint(scenes.WHEEL, 0, 1, local_axis_a=(0.0, 1.0), frequency_hz=1.5, damping_ratio=0.0)],
               gravity=(0, 0))
    o, st = run(sc, 400)
    y = np.array([s["cy"][1] for s in st])
    period = 2 * np.mean(np.diff(half_periods(y))) * DT
    assert 1 / 1.5 < period < 1.05 / 1.5
    assert np.abs([s["cx"][1] for s in st]).max() < 1e-5  # no motion across the axis


def test_friction_joint_decelerates_at_max_force_over_mass():
    sc = world([(0, 0, 0, 0), (2, 0, 0, 0)], [box(1, 0.5, 0.5)],
               [joint(scenes.FRICTION, 0, 1, max_force=2.0, max_torque=0.0)], gravity=(0, 0))
    sc.bodies["vx"][1] = 3.0
    o, st = run(sc, 120)
    m = 1.0
    for n in (1, 30, 60):
        assert st[n - 1]["vx"][1] == pytest.approx(3.0 - n * DT * 2.0 / m, rel=1e-4)
    assert abs(st[-1]["vx"][1]) < 1e-5  # stopped after 1.5 s and stays stopped


def test_motor_joint_reaches_its_offset():
    sc = world([(0, 0, 0, 0), (2, 0, 0, 0)], [box(1, 0.5, 0.5)],
               [joint(scenes.MOTOR, 0, 1, linear_offset=(2.0, 1.0), angular_offset=0.5, max_force=50.0,
                      max_torque=20.0, correction_factor=0.3)], gravity=(0, 0))
    o, st = run(sc, 600)
    s = st[-1]
    assert s["cx"][1] == pytest.approx(2.0, abs=1e-3) and s["cy"][1] == pytest.approx(1.0, abs=1e-3)
    assert s["a"][1] == pytest.approx(0.5, abs=1e-3)


def test_mouse_joint_pulls_the_body_to_the_target():
    sc = world([(0, 0, 0, 0), (2, 0, 5, 0)], [box(1, 0.5, 0.5)],
               [joint(scenes.MOUSE, 0, 1, target=(3.0, 5.0), max_force=1000.0)], gravity=(0, -10))
    o, st = run(sc, 240)
    s = st[-1]
    # the anchor is the target in body B's frame at creation, (3, 0): the body swings down
    # under it and the anchor sags below the target by about m g / k (k = m (2 pi 5 Hz)^2)
    ax = s["px"][1] + s["qc"][1] * 3.0
    ay = s["py"][1] + s["qs"][1] * 3.0
    assert math.hypot(ax - 3.0, ay - 5.0) < 0.05
    # without the joint the body would have fallen 80 m
    assert s["cy"][1] > 3.0


def test_pulley_conserves_rope_length():
    ga, gb = (-2.0, 10.0), (2.0, 10.0)
    pa, pb = (-2.0, 5.0), (2.0, 6.0)
    la, lb = 5.0, 4.0
    sc = world([(0, 0, 0, 0), (2, pa[0], pa[1], 0), (2, pb[0], pb[1], 0)], [box(1, 0.5, 0.5, 2.0), box(2, 0.5, 0.5, 1.0)],
               [joint(scenes.PULLEY, 1, 2, ground_anchor_a=ga, ground_anchor_b=gb, length_a=la, length_b=lb, ratio=1.0)])
    o, st = run(sc, 60)
    for s in st:
        a = math.hypot(s["cx"][1] - ga[0], s["cy"][1] - ga[1])
        b = math.hypot(s["cx"][2] - gb[0], s["cy"][2] - gb[1])
        assert abs(a + b - (la + lb)) < 0.01
    assert st[-1]["cy"][1] < pa[1] - 0.5 and st[-1]["cy"][2] > pb[1] + 0.5  # the 2 kg side goes down


# ---------------------------------------------------------------- filtering and islands
@pytest.mark.parametrize("collide", [0, 1])
def test_collide_connected_decides_the_contact(collide):
    sc = world([(2, 0, 0, 0), (2, 0.5, 0, 0)], [box(0, 0.5, 0.5), box(1, 0.5, 0.5)],
               [joint(scenes.REVOLUTE, 0, 1, local_anchor_a=(0.25, 0), local_anchor_b=(-0.25, 0),
                      collide_connected=collide)], gravity=(0, 0))
    o = po.OracleWorld(sc)
    o.step(DT)
    assert len(o.contacts()) == collide


def test_jointed_bodies_share_an_island():
    # two boxes resting apart on the ground joined by a rope: they fall asleep at the
    # same step, and a push on one wakes the other through the joint
    ground = scenes.fixture_defs(1)[0]
    scenes._edge(ground, 0, -20, 0, 20, 0, 0.6)
    b = scenes.body_defs(3)
    b[0]["type"] = 0
    b[1]["type"] = b[2]["type"] = 2
    b[1]["px"], b[1]["py"] = -3.0, 0.52
    b[2]["px"], b[2]["py"] = 3.0, 0.52
    b[2]["vx"] = 0.3  # keeps moving a while after body 1 has stopped
    f = np.array([ground, box(1, 0.5, 0.5), box(2, 0.5, 0.5, friction=0.05)], scenes.FIXTURE_DTYPE)
    j = np.array([joint(scenes.ROPE, 1, 2, length=10.0)], scenes.JOINT_DTYPE)
    o = po.OracleWorld(scenes.Scene(b, f, joints=j))
    slept = [None, None]
    for n in range(400):
        o.step(DT)
        s = o.bodies()
        for k in (1, 2):
            if slept[k - 1] is None and not (s["flags"][k] & 2):
                slept[k - 1] = n
    assert slept[0] is not None and slept[0] == slept[1]
    # without the joint body 1 sleeps earlier than body 2
    o2 = po.OracleWorld(scenes.Scene(b, f))
    first = [None, None]
    for n in range(400):
        o2.step(DT)
        s = o2.bodies()
        for k in (1, 2):
            if first[k - 1] is None and not (s["flags"][k] & 2):
                first[k - 1] = n
    assert first[0] < first[1]


# ---------------------------------------------------------------- gear
def _gear_wheels(ratio, w0, ra=0.5, rb=0.8):
    # wheel A (body 1) and wheel B (body 2) on revolute joints at their centres to the
    # static ground; joint 2 gears them (b2GearJoint: coordinate A + ratio * coordinate B)
    sc = world([(0, 0, 0, 0), (2, 0, 0, 0), (2, 3, 0, 0)], [circle(1, ra), circle(2, rb)],
               [joint(scenes.REVOLUTE, 0, 1), joint(scenes.REVOLUTE, 0, 2, local_anchor_a=(3.0, 0.0)),
                joint(scenes.GEAR, 1, 2, joint1=0, joint2=1, ratio=ratio)], gravity=(0, 0))
    sc.bodies["w"][1] = w0
    return sc


@pytest.mark.parametrize("ratio", [2.0, -0.5])
def test_gear_wheels_closed_form(ratio):
    # one velocity constraint wA + ratio * wB = 0: the first step's impulse is
    # P = -w0 / (iA + ratio^2 iB), leaving wA = w0 ratio^2 iB / (iA + ratio^2 iB) and
    # wB = -ratio iB w0 / (iA + ratio^2 iB), after which nothing changes; the angles keep
    # aA + ratio * aB = 0 (its value when the gear was made)
    ra, rb, w0 = 0.5, 0.8, 3.0
    iA = 1.0 / (0.5 * math.pi * ra ** 4)
    iB = 1.0 / (0.5 * math.pi * rb ** 4)
    den = iA + ratio * ratio * iB
    # one velocity iteration: the impulse is -Cdot / mass exactly, which pins the effective mass
    o1 = po.OracleWorld(_gear_wheels(ratio, w0, ra, rb), allow_sleep=False)
    o1.step(DT, 1, 3)
    assert o1.bodies()["w"][1] == pytest.approx(w0 * ratio * ratio * iB / den, rel=2e-5)
    o, st = run(_gear_wheels(ratio, w0, ra, rb), 120)
    for s in st:
        assert s["w"][1] == pytest.approx(w0 * ratio * ratio * iB / den, rel=2e-5)
        assert s["w"][2] == pytest.approx(-ratio * iB * w0 / den, rel=2e-5)
        assert abs(s["a"][1] + ratio * s["a"][2]) < 1e-4
    assert abs(st[-1]["a"][1]) > 0.2  # it turned
    # the wheels stay on their pivots
    assert np.abs(st[-1]["cx"][1:] - (0.0, 3.0)).max() < 1e-4


@pytest.mark.parametrize("rack_first", [True, False])
def test_gear_rack_and_pinion_closed_form(rack_first):
    # a rack (body 1) on a prismatic joint along x and a pinion (body 2) on a revolute
    # joint.  Rack as joint1, ratio r: vx + r w = 0 after the first impulse
    # P = -v0 / (m + r^2 i), vx = v0 r^2 i / (m + r^2 i), w = -r i v0 / (m + r^2 i).
    # Pinion as joint1: w + r vx = 0, P = -v0 r / (i + r^2 m) for the same push,
    # vx = v0 + r m P, w = i P (the prismatic Jacobian then carries the ratio)
    r, v0, rb = 0.5, 2.0, 0.5
    prism = joint(scenes.PRISMATIC, 0, 1, local_anchor_a=(0.0, 1.0), local_axis_a=(1.0, 0.0))
    rev = joint(scenes.REVOLUTE, 0, 2)
    js = [prism, rev] if rack_first else [rev, prism]
    gear = joint(scenes.GEAR, 1, 2, joint1=0, joint2=1, ratio=r) if rack_first else \
        joint(scenes.GEAR, 2, 1, joint1=0, joint2=1, ratio=r)
    sc = world([(0, 0, 0, 0), (2, 0, 1, 0), (2, 0, 0, 0)], [box(1, 2.0, 0.25), circle(2, rb)], js + [gear],
               gravity=(0, 0))
    sc.bodies["vx"][1] = v0
    m = 1.0 / (4.0 * 0.5)
    i = 1.0 / (0.5 * math.pi * rb ** 4)
    if rack_first:
        vx, w = v0 * r * r * i / (m + r * r * i), -r * i * v0 / (m + r * r * i)
    else:
        P = -v0 * r / (i + r * r * m)
        vx, w = v0 + r * m * P, i * P
    o1 = po.OracleWorld(sc, allow_sleep=False)
    o1.step(DT, 1, 3)  # one iteration: pins the effective mass
    assert o1.bodies()["vx"][1] == pytest.approx(vx, rel=2e-5)
    o, st = run(sc, 90)
    for s in st:
        assert s["vx"][1] == pytest.approx(vx, rel=2e-5)
        assert s["w"][2] == pytest.approx(w, rel=2e-5)
        lin, ang = s["cx"][1], s["a"][2]
        assert abs((lin + r * ang) if rack_first else (ang + r * lin)) < 1e-4
        assert abs(s["vy"][1]) < 1e-6 and abs(s["w"][1]) < 1e-6


def test_gear_reacts_on_a_dynamic_carrier():
    # wheel A turns on a revolute joint of a dynamic carrier disk C, coaxial with it,
    # and C turns on a revolute joint of the ground at the same point; wheel B turns on
    # the ground (D static).  Every body only rotates about its pin, so the first
    # impulse of the velocity constraint (wA - wC) + r wB = 0 is
    # P = -w0 / (iA + iC + r^2 iB): wA = w0 + iA P, wC = -iC P (the carrier takes the
    # reaction), wB = r iB P; a sign or inertia error on C moves all three
    r, w0, ra, rc, rb = 1.5, 3.0, 0.4, 0.7, 0.6
    sc = world([(0, 0, 0, 0), (2, 0, 0, 0), (2, 0, 0, 0), (2, 3, 0, 0)],
               [circle(1, ra), circle(2, rc), circle(3, rb)],
               [joint(scenes.REVOLUTE, 0, 2), joint(scenes.REVOLUTE, 2, 1), joint(scenes.REVOLUTE, 0, 3,
                                                                                   local_anchor_a=(3.0, 0.0)),
                joint(scenes.GEAR, 1, 3, joint1=1, joint2=2, ratio=r)], gravity=(0, 0))
    sc.bodies["w"][1] = w0
    inv = lambda rad: 1.0 / (0.5 * math.pi * rad ** 4)  # noqa: E731
    iA, iC, iB = inv(ra), inv(rc), inv(rb)
    P = -w0 / (iA + iC + r * r * iB)
    o1 = po.OracleWorld(sc, allow_sleep=False)
    o1.step(DT, 1, 3)  # one iteration: pins the effective mass
    assert o1.bodies()["w"][2] == pytest.approx(-iC * P, rel=2e-5)
    o, st = run(sc, 60)
    for s_ in st:
        assert s_["w"][1] == pytest.approx(w0 + iA * P, rel=2e-5)
        assert s_["w"][2] == pytest.approx(-iC * P, rel=2e-5)
        assert s_["w"][3] == pytest.approx(r * iB * P, rel=2e-5)
        assert abs((s_["a"][1] - s_["a"][2]) + r * s_["a"][3]) < 1e-4
    # the gear joint's impulse is its first read-back value: the first step's P,
    # warm started and removed again on alternate steps
    assert o.joints()[3][0] in (pytest.approx(0.0, abs=1e-6), pytest.approx(P, rel=1e-5))


def test_gear_definitions_checked():
    # the gear couples revolute or prismatic joints made before it
    base = [(0, 0, 0, 0), (2, 0, 0, 0), (2, 3, 0, 0)]
    for bad in ([joint(scenes.DISTANCE, 0, 1), joint(scenes.REVOLUTE, 0, 2),
                 joint(scenes.GEAR, 1, 2, joint1=0, joint2=1, ratio=1.0)],
                [joint(scenes.REVOLUTE, 0, 1), joint(scenes.GEAR, 1, 2, joint1=0, joint2=2, ratio=1.0),
                 joint(scenes.REVOLUTE, 0, 2)]):
        sc = world(base, [circle(1, 0.5), circle(2, 0.5)], bad, gravity=(0, 0))
        with pytest.raises(Exception):
            po.OracleWorld(sc)
