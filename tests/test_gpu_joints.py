"""GPU joints against the oracle, bit for bit (the oracle's b2Rot.Set evaluated as the
kernels evaluate it, exact_rot=True): every joint kind the library supports, alone
in the closed-form scenes that pin the oracle (tests/test_oracle_joints.py) and all
together with contacts, continuous physics and sleeping in the joint world; joint
worlds batched equal the same worlds alone; definitions the library rejects."""
from __future__ import annotations

import numpy as np
import pytest

from dbox_amd import native, scenes
from oracle import pyoracle as po
from tests import test_oracle_joints as tj
from tests.helpers import body_mismatch, contact_mismatch

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0
DEFAULTS = native.WARM | native.SLEEP | native.CCD | native.ORDERED


def _gpu(scene_list):
    h = native.StepWorlds()
    for s in scene_list:
        h.add_scene(s)
    h.finalize()
    return h


def _lockstep(scene, steps, every=1, flags=DEFAULTS, allow_sleep=True, label=""):
    o = po.OracleWorld(scene, continuous=bool(flags & native.CCD), allow_sleep=allow_sleep, exact_rot=True)
    g = _gpu([scene])
    gflags = flags if allow_sleep else flags & ~native.SLEEP
    for i in range(steps):
        o.step(DT)
        g.step(DT, flags=gflags)
        if (i + 1) % every == 0 or i + 1 == steps:
            m = body_mismatch(g.bodies(), o.bodies(), o.bodies_def["type"])
            assert m is None, f"{label} step {i + 1}: {m}"
            gj, oj = g.joints(), o.joints()
            assert gj.tobytes() == oj.tobytes(), f"{label} step {i + 1}: joints {gj} vs {oj}"
            m = contact_mismatch(g.contacts(), o.contacts())
            assert m is None, f"{label} step {i + 1}: {m}"
    return o, g


def _closed_form_scenes():
    W, J, S = tj.world, tj.joint, scenes
    c, b = tj.circle, tj.box
    import math
    th = 0.08
    yield "distance pendulum", W([(0, 0, 10, 0), (2, 2 * math.sin(th), 10 - 2 * math.cos(th), 0)], [c(1, 0.05)],
                                 [J(S.DISTANCE, 0, 1, length=2.0)])
    yield "soft distance", W([(0, 0, 0, 0), (2, 1.3, 0, 0)], [c(1, 0.56)],
                             [J(S.DISTANCE, 0, 1, length=1.0, frequency_hz=2.0, damping_ratio=0.0)], gravity=(0, 0))
    yield "revolute pendulum", W([(0, 0, 10, 0), (2, 0.087, 9.004, 0.087 - math.pi / 2)], [b(1, 1.0, 0.1)],
                                 [J(S.REVOLUTE, 0, 1, local_anchor_b=(-1.0, 0.0))])
    yield "revolute motor", W([(0, 0, 0, 0), (2, 0, 0, 0)], [c(1, 0.5)],
                              [J(S.REVOLUTE, 0, 1, enable_motor=1, motor_speed=100.0, max_motor=0.05)], gravity=(0, 0))
    yield "revolute limit", W([(0, 0, 10, 0), (2, 1.0, 10, 0)], [b(1, 1.0, 0.1)],
                              [J(S.REVOLUTE, 0, 1, local_anchor_b=(-1.0, 0.0), enable_limit=1, lower=-0.5, upper=0.5)])
    yield "prismatic slope", W([(0, 0, 0, 0), (2, 0, 0, 0)], [b(1, 0.5, 0.5)],
                               [J(S.PRISMATIC, 0, 1, local_axis_a=(0.866, -0.5))])
    yield "prismatic limit", W([(0, 0, 0, 0), (2, 0, 0, 0)], [b(1, 0.5, 0.5)],
                               [J(S.PRISMATIC, 0, 1, local_axis_a=(0.0, 1.0), enable_limit=1, lower=-1.0, upper=1.0)])
    g = scenes.fixture_defs(1)[0]
    scenes._edge(g, 0, -20, 0, 20, 0, 0.6)
    yield "weld", W([(0, 0, 0, 0), (2, 0, 3, 0.3), (2, 1.2, 3.4, 0.3)], [g, b(1, 0.5, 0.2), b(2, 0.5, 0.2)],
                    [J(S.WELD, 1, 2, local_anchor_a=(0.6, 0.2), local_anchor_b=(-0.6, -0.2))])
    yield "rope", W([(0, 0, 10, 0), (2, 0, 9, 0)], [c(1, 0.1)], [J(S.ROPE, 0, 1, length=3.0)])
    yield "wheel", W([(0, 0, 0, 0), (2, 0, 0.3, 0)], [c(1, 0.56)],
                     [J(S.WHEEL, 0, 1, local_axis_a=(0.0, 1.0), frequency_hz=1.5, damping_ratio=0.0)], gravity=(0, 0))
    fr = W([(0, 0, 0, 0), (2, 0, 0, 0)], [b(1, 0.5, 0.5)], [J(S.FRICTION, 0, 1, max_force=2.0, max_torque=0.0)],
           gravity=(0, 0))
    fr.bodies["vx"][1] = 3.0
    yield "friction", fr
    yield "motor", W([(0, 0, 0, 0), (2, 0, 0, 0)], [b(1, 0.5, 0.5)],
                     [J(S.MOTOR, 0, 1, linear_offset=(2.0, 1.0), angular_offset=0.5, max_force=50.0, max_torque=20.0,
                        correction_factor=0.3)], gravity=(0, 0))
    yield "mouse", W([(0, 0, 0, 0), (2, 0, 5, 0)], [b(1, 0.5, 0.5)],
                     [J(S.MOUSE, 0, 1, target=(3.0, 5.0), max_force=1000.0)])
    yield "pulley", W([(0, 0, 0, 0), (2, -2, 5, 0), (2, 2, 6, 0)], [b(1, 0.5, 0.5, 2.0), b(2, 0.5, 0.5, 1.0)],
                      [J(S.PULLEY, 1, 2, ground_anchor_a=(-2, 10), ground_anchor_b=(2, 10), length_a=5.0, length_b=4.0,
                         ratio=1.0)])
    gw = tj._gear_wheels(2.0, 3.0)
    yield "gear wheels", gw
    for first in (True, False):
        prism = J(S.PRISMATIC, 0, 1, local_anchor_a=(0.0, 1.0), local_axis_a=(0.8, 0.6))
        rev = J(S.REVOLUTE, 0, 2, local_anchor_a=(0.0, 0.0))
        rp = W([(0, 0, 0, 0), (2, 0, 1, 0.2), (2, 0, 0, 0.7)], [b(1, 2.0, 0.25), c(2, 0.5)],
               ([prism, rev] if first else [rev, prism]) + [J(S.GEAR, 1, 2, joint1=0, joint2=1, ratio=0.5)])
        rp.bodies["vx"][1] = 2.0
        yield f"rack and pinion {'rack' if first else 'pinion'} first", rp
    # wheels on revolute joints of one falling dynamic frame: the gear's C and D are the
    # same dynamic body, whose velocity the reference writes twice (C, then D)
    fr = W([(2, 0, 5, 0), (2, -1, 5, 0), (2, 1, 5, 0), (0, 0, 0, 0)],
           [b(0, 1.5, 0.2), c(1, 0.4), c(2, 0.6), g],
           [J(S.REVOLUTE, 0, 1, local_anchor_a=(-1.0, 0.0)), J(S.REVOLUTE, 0, 2, local_anchor_a=(1.0, 0.0)),
            J(S.GEAR, 1, 2, joint1=0, joint2=1, ratio=-1.5)])
    fr.fixtures["body"][3] = 3
    fr.bodies["w"][1] = 6.0
    yield "gear on a shared frame", fr


@pytest.mark.parametrize("name,scene", list(_closed_form_scenes()), ids=[n for n, _ in _closed_form_scenes()])
def test_closed_form_joint_scene_bit_identical(name, scene):
    print(f"case: joint scene '{name}', 240 steps, exact sin/cos", flush=True)
    _lockstep(scene, 240, every=1, allow_sleep=False, label=name)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_joint_world_bit_identical(seed):
    # every joint kind beside contacts, continuous physics and sleeping, 400 steps
    print(f"case: joint world seed {seed}, 400 steps, dbox defaults, exact sin/cos", flush=True)
    o, g = _lockstep(scenes.joint_world(seed), 400, every=5, label=f"joint world {seed}")
    print(f"joint world {seed}: {len(o.contacts())} contacts, awake {int(((o.bodies()['flags'] & 2) != 0).sum())}")


def test_joint_worlds_batched_equal_single_and_coloured_mode():
    # joint worlds beside plain container worlds in one batch; worlds with joints are
    # solved in the exact order whatever the mode, so the coloured flag gives the same bits
    print("case: 24 joint worlds + 24 containers batched vs single, 200 steps", flush=True)
    ws = []
    for k in range(24):
        ws.append(scenes.joint_world(100 + k))
        ws.append(scenes.mixed_container(60 + 5 * k, 40.0, 80.0, seed=900 + k))
    g = _gpu(ws)
    g.step(flags=DEFAULTS, n=200)
    c = _gpu(ws)
    c.step(flags=DEFAULTS & ~native.ORDERED, n=200)
    for k in (0, 10, 46):
        single = _gpu([ws[k]])
        single.step(flags=DEFAULTS, n=200)
        assert g.bodies(k).tobytes() == single.bodies(0).tobytes(), k
        assert g.contacts(k).tobytes() == single.contacts(0).tobytes(), k
        if k % 2 == 0:
            assert g.joints(k).tobytes() == single.joints(0).tobytes(), k
            assert c.bodies(k).tobytes() == g.bodies(k).tobytes(), k
        single.close()


@pytest.mark.parametrize("collide", [0, 1])
def test_collide_connected_on_the_gpu(collide):
    print(f"case: collideConnected {collide}", flush=True)
    sc = tj.world([(2, 0, 0, 0), (2, 0.5, 0, 0)], [tj.box(0, 0.5, 0.5), tj.box(1, 0.5, 0.5)],
                  [tj.joint(scenes.REVOLUTE, 0, 1, local_anchor_a=(0.25, 0), local_anchor_b=(-0.25, 0),
                            collide_connected=collide)], gravity=(0, 0))
    g = _gpu([sc])
    g.step()
    assert len(g.contacts()) == collide


def test_joint_definitions_rejected():
    print("case: rejected joint definitions", flush=True)
    sc = tj.world([(0, 0, 0, 0), (2, 0, 1, 0)], [tj.box(1, 0.5, 0.5)], [])
    h = native.StepWorlds()
    w = h.add_scene(sc)
    # a gear must couple revolute or prismatic joints made before it; a failed batch adds nothing
    pair = np.concatenate([scenes.joint_defs(1, scenes.DISTANCE), scenes.joint_defs(1, scenes.GEAR)])
    pair["body_a"], pair["body_b"] = 0, 1
    pair["joint1"][1], pair["joint2"][1] = 0, 0  # joint 0 is a distance joint
    assert h.L.dbx_add_joints(h.h, w, native._p(pair), 2) == native.DBX_E_INVALID_ARG
    gear = scenes.joint_defs(1, scenes.GEAR)
    gear["body_a"], gear["body_b"], gear["joint1"], gear["joint2"] = 0, 1, 0, 1  # no joints yet
    assert h.L.dbx_add_joints(h.h, w, native._p(gear), 1) == native.DBX_E_INVALID_ARG
    bad = scenes.joint_defs(1, scenes.REVOLUTE)
    bad["body_a"], bad["body_b"] = 1, 1
    assert h.L.dbx_add_joints(h.h, w, native._p(bad), 1) == native.DBX_E_INVALID_ARG
    nan = scenes.joint_defs(1, scenes.DISTANCE)
    nan["body_a"], nan["body_b"], nan["length"] = 0, 1, np.nan
    assert h.L.dbx_add_joints(h.h, w, native._p(nan), 1) == native.DBX_E_INVALID_ARG
    h.finalize()
    assert h.L.dbx_read_joints(h.h, w, None, 0) == native.DBX_OK  # none of the rejected joints was added
    ok = scenes.joint_defs(1, scenes.DISTANCE)
    ok["body_a"], ok["body_b"] = 0, 1
    assert h.L.dbx_add_joints(h.h, w, native._p(ok), 1) == native.DBX_E_LOCKED
