"""The oracle's sensor fixtures (b2Fixture.m_isSensor; b2Contact.Update's sensor
branch, b2contact.d:289-298, with b2TestOverlap, b2collision.d:449-468), checked
against what fixes them without running dbox:

* a circle falling through a static circle sensor: BEGIN at the first Collide
  after the centres come within the sum of the radii, END at the first after they
  part, and the fall is bit-identical to the same world without the sensor (no
  manifold, no impulse, no PostSolve, no PreSolve);
* a box falling through a box sensor overlaps while the vertical gap between the
  cores is below the two polygon skins (2 x b2_polygonRadius);
* a kinematic sensor sweeping over sleeping boxes reports them without waking
  them (b2contactmanager.d:169-173 and the sensor branch wake nothing).
"""
from __future__ import annotations

import numpy as np

from dbox_amd import scenes
from oracle import pyoracle as po

SKIN = 2 * 0.005  # b2_polygonRadius


def _scene(bodies, fixtures):
    b = scenes.body_defs(len(bodies))
    for i, (t, x, y) in enumerate(bodies):
        b[i]["type"], b[i]["px"], b[i]["py"] = t, x, y
    return scenes.Scene(b, np.array(fixtures, scenes.FIXTURE_DTYPE))


def _circle(body, r, sensor=False, density=1.0):
    f = scenes.fixture_defs(1)[0]
    f["body"], f["shape"], f["radius"], f["density"] = body, scenes.SHAPE_CIRCLE, r, density
    f["flags"] = scenes.SENSOR if sensor else 0
    return f


def _box(body, hx, hy, sensor=False, density=1.0):
    f = scenes.fixture_defs(1)[0]
    f["body"], f["shape"], f["density"] = body, scenes.SHAPE_BOX, density
    f["p"][:2] = (hx, hy)
    f["flags"] = scenes.SENSOR if sensor else 0
    return f


def test_circle_through_circle_sensor_closed_form():
    R, r = 2.0, 0.5
    with_sensor = _scene([(scenes.STATIC, 0.0, 0.0), (scenes.DYNAMIC, 0.3, 6.0)],
                         [_circle(0, R, sensor=True), _circle(1, r)])
    without = _scene([(scenes.STATIC, 0.0, 0.0), (scenes.DYNAMIC, 0.3, 6.0)], [_circle(1, r)])
    a, b = po.OracleWorld(with_sensor, allow_sleep=False), po.OracleWorld(without, allow_sleep=False)
    a.set_presolve(1)
    prev = None
    begin = end = None
    for step in range(1, 151):
        a.step()
        b.step()
        assert a.bodies()[1].tobytes() == b.bodies()[1].tobytes(), step
        ev = a.events()
        assert not np.any(ev["type"] >= 3), step  # no PostSolve, no PreSolve
        kinds = list(ev["type"])
        if prev is not None:
            # Collide at the start of this step sees the positions the last step left
            dist = float(np.hypot(prev[0] - 0.0, prev[1] - 0.0))
            inside = dist < R + r
            if kinds == [1]:
                begin = step
                assert inside, (step, dist)
            elif kinds == [2]:
                end = step
                assert not inside, (step, dist)
            else:
                assert kinds == [], (step, kinds)
        c = a.contacts()
        if len(c):
            assert np.all(c["point_count"] == 0)
        bo = a.bodies()[1]
        prev = (float(bo["cx"]), float(bo["cy"]))
        if begin is not None and end is None:
            assert (c["flags"] & 2).any()  # touching while inside
    assert begin is not None and end is not None and end > begin


def test_box_through_box_sensor_uses_the_skins():
    hx, hy, h = 3.0, 0.2, 0.25
    sc = _scene([(scenes.STATIC, 0.0, 0.0), (scenes.DYNAMIC, 0.0, 3.0)],
                [_box(0, hx, hy, sensor=True), _box(1, h, h)])
    o = po.OracleWorld(sc, allow_sleep=False)
    prev_gap = None
    seen = []
    for step in range(1, 80):
        o.step()
        kinds = list(o.events()["type"])
        if prev_gap is not None and kinds:
            seen.append((kinds, prev_gap))
        y = float(o.bodies()[1]["cy"])
        # vertical gap between the cores while the box is above the sensor, below it after
        prev_gap = y - h - hy if y > 0 else -(y + h) - hy
    (k1, g1), (k2, g2) = seen
    assert k1 == [1] and g1 < SKIN and k2 == [2] and g2 >= SKIN, seen


def test_kinematic_sensor_does_not_wake_sleepers():
    boxes = [_box(1 + i, 0.5, 0.5) for i in range(4)]
    ground = scenes.fixture_defs(1)[0]
    ground["body"], ground["shape"] = 0, scenes.SHAPE_EDGE
    ground["p"][:4] = (-20.0, 0.0, 20.0, 0.0)
    sweeper = _box(5, 1.0, 3.0, sensor=True)
    bodies = [(scenes.STATIC, 0.0, 0.0)] + [(scenes.DYNAMIC, -6.0 + 4.0 * i, 0.5) for i in range(4)]
    bodies.append((scenes.KINEMATIC, -40.0, 1.0))
    sc = _scene(bodies, [ground] + boxes + [sweeper])
    sc.bodies["vx"][5] = 12.0  # 0.2 m per step: it reaches the first box after 160 steps
    o = po.OracleWorld(sc)
    o.step(n=100)  # the boxes fall asleep
    assert not np.any(o.bodies()["flags"][1:5] & 2)
    begins = ends = 0
    for _ in range(250):
        o.step()
        ev = o.events()
        begins += int(np.count_nonzero(ev["type"] == 1))
        ends += int(np.count_nonzero(ev["type"] == 2))
        assert not np.any(o.bodies()["flags"][1:5] & 2)  # still asleep
    assert begins == 4 and ends == 4
