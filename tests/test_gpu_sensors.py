"""GPU parity of sensor fixtures (b2Fixture.m_isSensor): b2Contact.Update's sensor
branch (b2contact.d:289-298) with b2TestOverlap (b2collision.d:449-468), no wake
at creation (b2contactmanager.d:169-173), no island, no TOI (b2world.d:1006,
1176-1181, 1311-1316).

* the sensor container (a static sensor box across the middle, every 9th body a
  sensor that falls through everything), dbox's defaults with continuous physics:
  every step's contacts (pairs, order, touching flags, point counts) and callbacks
  (kinds, pairs, order, counts) equal the oracle's, while the runs agree (bodies
  within 1e-4 m, impulses within 1e-3 relative / 2e-3 absolute), and sensor
  contacts have no points;
* the same world with split steps: no PRE_SOLVE for a sensor contact;
* a circle through a circle sensor and a box through a box sensor: BEGIN / END
  on the same steps as the oracle, the fall identical to the world without it;
* a kinematic sensor sweeping over sleeping boxes reports all four, waking none.
"""
from __future__ import annotations

import numpy as np
import pytest

from dbox_amd import native, scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
FLAGS = native.WARM | native.SLEEP | native.ORDERED | native.CCD


def _gpu(scene_list):
    h = native.StepWorlds()
    for s in scene_list:
        h.add_scene(s)
    h.finalize()
    return h


def _lockstep(scene, steps, presolve=False, tol=1e-4):
    o = po.OracleWorld(scene, continuous=True)
    if presolve:
        o.set_presolve(1)
    g = _gpu([scene])
    compared = sensed = 0
    for i in range(steps):
        o.step()
        g.step(flags=FLAGS | (native.PRESOLVE if presolve else 0))
        if presolve:
            g.step_finish()
        bo, bg = o.bodies(), g.bodies()
        dp = max(np.abs(bo["cx"] - bg["cx"]).max(), np.abs(bo["cy"] - bg["cy"]).max())
        if dp > tol:
            print(f"states separate at step {i + 1} ({dp:.2e})")
            break
        co, cg = o.contacts(), g.contacts()
        assert len(co) == len(cg), (i, len(co), len(cg))
        for f in ("fixture_a", "fixture_b", "point_count"):
            assert np.array_equal(cg[f], co[f]), (i, f)
        assert np.array_equal(cg["flags"] & 6, co["flags"] & 6), i
        eo, eg = o.events(), g.events()
        assert len(eg) == len(eo), (i, len(eg), len(eo))
        for f in ("type", "fixture_a", "fixture_b", "point_count"):
            assert np.array_equal(eg[f], eo[f]), (i, f)
        # the solver's impulses drift apart (sin/cos rounding) before the positions do
        if not all(np.allclose(eg[f], eo[f], rtol=1e-3, atol=2e-3) for f in ("normal_impulse", "tangent_impulse")):
            print(f"impulses separate at step {i + 1}")
            break
        compared += 1
        sensed += int(np.count_nonzero(eg["type"] <= 2))
    return g, compared, sensed


def _sensor_proxies(scene):
    flags = np.repeat(scene.fixtures["flags"], [scenes.proxy_count(scene.fixtures[i:i + 1])
                                                for i in range(len(scene.fixtures))])
    return np.flatnonzero(flags & scenes.SENSOR)


def test_sensor_container_tracks_the_oracle():
    print("case: sensor container, 150 bodies, 120 steps", flush=True)
    s = scenes.sensor_container(150)
    g, compared, sensed = _lockstep(s, 120)
    print(f"contacts and callbacks equal for {compared} steps, {sensed} BEGIN/END")
    assert compared >= 60 and sensed > 50
    c = g.contacts()
    sens = np.isin(c["fixture_a"], _sensor_proxies(s)) | np.isin(c["fixture_b"], _sensor_proxies(s))
    assert sens.any() and np.all(c["point_count"][sens] == 0)
    assert np.count_nonzero((c["flags"][sens] & 2) != 0) > 0


def test_sensor_container_split_steps_give_no_sensor_presolve():
    print("case: sensor container, split steps, 80 steps", flush=True)
    s = scenes.sensor_container(150, seed=62)
    g, compared, sensed = _lockstep(s, 80, presolve=True)
    assert compared >= 40 and sensed > 20


def _two(shape_a, shape_b, y):
    b = scenes.body_defs(2)
    b[0]["type"] = scenes.STATIC
    b[1]["type"] = scenes.DYNAMIC
    b[1]["px"], b[1]["py"] = 0.3, y
    f = scenes.fixture_defs(2)
    for k, (shape, dims) in enumerate((shape_a, shape_b)):
        f[k]["body"] = k
        f[k]["shape"] = shape
        f[k]["density"] = 1.0
        if shape == scenes.SHAPE_CIRCLE:
            f[k]["radius"] = dims
        else:
            f[k]["p"][:2] = dims
    f[0]["flags"] = scenes.SENSOR
    return scenes.Scene(b, f)


def test_single_sensor_overlaps_like_the_oracle():
    print("case: circle and box through sensors, 150 steps", flush=True)
    for sc in (_two((scenes.SHAPE_CIRCLE, 2.0), (scenes.SHAPE_CIRCLE, 0.5), 6.0),
               _two((scenes.SHAPE_BOX, (3.0, 0.2)), (scenes.SHAPE_BOX, (0.25, 0.25)), 3.0)):
        o = po.OracleWorld(sc, continuous=True)
        g = _gpu([sc])
        kinds = []
        for i in range(150):
            o.step()
            g.step(flags=FLAGS)
            eo, eg = o.events(), g.events()
            assert np.array_equal(eg["type"], eo["type"]), i
            kinds += list(eg["type"])
            np.testing.assert_allclose(g.bodies()["cy"], o.bodies()["cy"], rtol=0, atol=1e-4)
        assert kinds == [native.EVENT_BEGIN, native.EVENT_END], kinds


def test_kinematic_sensor_wakes_no_sleeper():
    print("case: kinematic sensor over sleeping boxes, 350 steps", flush=True)
    b = scenes.body_defs(6)
    b["type"] = [scenes.STATIC] + [scenes.DYNAMIC] * 4 + [scenes.KINEMATIC]
    b["px"] = [0.0, -6.0, -2.0, 2.0, 6.0, -40.0]
    b["py"] = [0.0, 0.5, 0.5, 0.5, 0.5, 1.0]
    b[5]["vx"] = 12.0
    f = scenes.fixture_defs(6)
    f[0]["shape"] = scenes.SHAPE_EDGE
    f[0]["p"][:4] = (-20.0, 0.0, 20.0, 0.0)
    for k in range(1, 6):
        f[k]["body"] = k
        f[k]["shape"] = scenes.SHAPE_BOX
        f[k]["density"] = 1.0
        f[k]["p"][:2] = (0.5, 0.5) if k < 5 else (1.0, 3.0)
    f[5]["flags"] = scenes.SENSOR
    sc = scenes.Scene(b, f)
    g = _gpu([sc])
    g.step(flags=FLAGS, n=100)
    assert not np.any(g.bodies()["flags"][1:5] & 2)
    begins = ends = 0
    for _ in range(250):
        g.step(flags=FLAGS)
        ev = g.events()
        begins += int(np.count_nonzero(ev["type"] == native.EVENT_BEGIN))
        ends += int(np.count_nonzero(ev["type"] == native.EVENT_END))
        assert not np.any(g.bodies()["flags"][1:5] & 2)
    assert begins == 4 and ends == 4
