"""GPU parity of the continuous pass's listener callbacks (b2World.SolveTOI,
b2world.d:1127-1452): BEGIN / END from each b2Contact.Update of a TOI event
(b2contact.d:342-350), PRE_SOLVE after a split step (b2contact.d:352-355), and
POST_SOLVE for the sub-step island with the sub-step solver's impulses
(b2island.d:415,438-461), after the discrete step's callbacks.

* the fast box over a thin edge: every step's callbacks equal the oracle's (the
  landing step's BEGIN, then the TOI island's POST_SOLVE);
* the chain terrain and the projectile room: every step's callbacks equal the
  oracle's in kind, pair, order and point count, impulses within 1e-3 relative /
  2e-3 absolute, while the two runs agree;
* a split step with continuous physics reports the pass's PRE_SOLVE calls too.
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


def _fast_box():
    b = scenes.body_defs(2)
    b[0]["type"] = scenes.STATIC
    b[1]["type"] = scenes.DYNAMIC
    b[1]["py"] = 1.2
    b[1]["vy"] = -120.0
    f = scenes.fixture_defs(2)
    f[0]["body"] = 0
    f[0]["shape"] = scenes.SHAPE_EDGE
    f[0]["p"][:4] = (-5.0, 0.0, 5.0, 0.0)
    f[1]["body"] = 1
    f[1]["shape"] = scenes.SHAPE_BOX
    f[1]["p"][:2] = (0.25, 0.25)
    f[1]["density"] = 1.0
    return scenes.Scene(b, f, name="fast_box")


def _assert_same_events(eg, eo, step):
    assert len(eg) == len(eo), (step, len(eg), len(eo))
    for f in ("type", "fixture_a", "fixture_b", "point_count"):
        assert np.array_equal(eg[f], eo[f]), (step, f)
    for f in ("normal_impulse", "tangent_impulse"):
        np.testing.assert_allclose(eg[f], eo[f], rtol=1e-3, atol=2e-3, err_msg=f"step {step} {f}")


def _lockstep(scene, steps, flags=FLAGS, presolve=False):
    o = po.OracleWorld(scene, continuous=True)
    if presolve:
        o.set_presolve(1)
    g = _gpu([scene])
    compared = toi_steps = 0
    for i in range(steps):
        o.step()
        g.step(flags=flags | (native.PRESOLVE if presolve else 0))
        if presolve:
            g.step_finish()
        bo, bg = o.bodies(), g.bodies()
        dp = max(np.abs(bo["cx"] - bg["cx"]).max(), np.abs(bo["cy"] - bg["cy"]).max())
        if dp > 1e-4:
            print(f"states separate at step {i + 1} ({dp:.2e})")
            break
        assert int(g.toi_events()[0]) == o.toi_events(), i
        _assert_same_events(g.events(), o.events(), i + 1)
        compared += 1
        toi_steps += int(o.toi_events() > 0)
    print(f"callbacks equal for {compared} steps, {toi_steps} of them with TOI events")
    return compared, toi_steps


def test_fast_box_toi_callbacks():
    print("case: fast box over an edge, continuous callbacks", flush=True)
    compared, toi_steps = _lockstep(_fast_box(), 30)
    assert compared == 30 and toi_steps >= 1
    # the landing step on its own: BEGIN from the TOI event's update, then its island's PostSolve
    g = _gpu([_fast_box()])
    g.step(flags=FLAGS)
    ev = g.events()
    assert list(ev["type"]) == [native.EVENT_BEGIN, native.EVENT_POST_SOLVE]
    assert abs(float(ev["normal_impulse"][1].sum()) - 0.25 * 120.0) < 0.03


def test_chain_terrain_toi_callbacks():
    print("case: chain terrain, 60 bodies, 90 steps, continuous callbacks", flush=True)
    compared, toi_steps = _lockstep(scenes.chain_terrain(60), 90)
    assert compared >= 60 and toi_steps >= 10


def test_projectile_room_toi_callbacks():
    print("case: projectile room, 40 bodies, continuous callbacks", flush=True)
    compared, toi_steps = _lockstep(scenes.projectile_room(40), 40)
    assert compared >= 2 and toi_steps >= 2


def test_split_step_reports_toi_presolve():
    print("case: chain terrain, split steps, continuous PreSolve reports", flush=True)
    compared, toi_steps = _lockstep(scenes.chain_terrain(60, seed=5), 60, presolve=True)
    assert compared >= 40 and toi_steps >= 5
