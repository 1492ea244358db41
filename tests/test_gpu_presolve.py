"""GPU parity of the PreSolve split: dbx_step with DBX_STEP_PRESOLVE stops after
b2ContactManager.Collide, the host reads the Collide callbacks (BEGIN / END /
PRE_SOLVE, b2contact.d:342-355), applies what its PreSolve changed with
dbx_write_contact_settings (b2Contact.SetEnabled / SetFriction / SetRestitution /
SetTangentSpeed, b2contact.d:108-198), and dbx_step_finish runs the rest.

* with the oracle's fixed test policy (by (fixtureA + fixtureB) % 7: disable,
  friction 0.05, restitution 0.6, tangent speed 0.8) applied from Python between
  the halves, every step's callbacks match the oracle's in kind, pair and order
  and the bodies track the oracle's within 1e-4 m while they agree;
* a split step with no settings gives the same bits as a plain step;
* the split's state errors.
"""
from __future__ import annotations

import numpy as np
import pytest

from dbox_amd import native, scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
FLAGS = native.WARM | native.SLEEP | native.ORDERED


def _gpu(scene_list):
    h = native.StepWorlds()
    for s in scene_list:
        h.add_scene(s)
    h.finalize()
    return h


def _policy(events):
    """The oracle's test PreSolve policy as contact settings."""
    pre = events[events["type"] == native.EVENT_PRE_SOLVE]
    k = (pre["fixture_a"] + pre["fixture_b"]) % 7
    sel = k <= 3
    s = np.zeros(int(sel.sum()), native.SETTING_DTYPE)
    s["fixture_a"] = pre["fixture_a"][sel]
    s["fixture_b"] = pre["fixture_b"][sel]
    s["mask"] = np.choose(k[sel], [native.SET_ENABLED, native.SET_FRICTION, native.SET_RESTITUTION,
                                   native.SET_TANGENT_SPEED])
    s["enabled"] = 0
    s["friction"] = 0.05
    s["restitution"] = 0.6
    s["tangent_speed"] = 0.8
    return s


def test_presolve_split_with_policy_matches_oracle():
    print("case: PreSolve split, 150-body container, policy, 120 steps", flush=True)
    sc = scenes.mixed_container(150, 40.0, 80.0, seed=32)
    o = po.OracleWorld(sc)
    o.set_presolve(2)
    g = _gpu([sc])
    compared = changed = 0
    for i in range(120):
        o.step()
        g.step(flags=FLAGS | native.PRESOLVE)
        mid = g.events()
        assert not np.any(mid["type"] == native.EVENT_POST_SOLVE)
        s = _policy(mid)
        changed += len(s)
        g.write_contact_settings(0, s)
        g.step_finish()
        bo, bg = o.bodies(), g.bodies()
        if max(np.abs(bo["cx"] - bg["cx"]).max(), np.abs(bo["cy"] - bg["cy"]).max()) > 1e-4:
            print(f"states separate at step {i + 1}")
            break
        eo, eg = o.events(), g.events()
        assert len(eg) == len(eo), (i, len(eg), len(eo))
        for f in ("type", "fixture_a", "fixture_b", "point_count"):
            assert np.array_equal(eg[f], eo[f]), (i, f)
        for f in ("normal_impulse", "tangent_impulse"):
            np.testing.assert_allclose(eg[f], eo[f], rtol=1e-3, atol=2e-3)
        compared += 1
    print(f"compared {compared} steps, {changed} contact settings written")
    assert compared >= 60 and changed > 500


def test_presolve_split_without_settings_equals_plain_step():
    print("case: PreSolve split without settings vs plain, 80 steps", flush=True)
    sc = scenes.mixed_container(150, 40.0, 80.0, seed=33)
    a, b = _gpu([sc]), _gpu([sc])
    for i in range(80):
        a.step(flags=FLAGS)
        b.step(flags=FLAGS | native.PRESOLVE)
        b.step_finish()
        assert a.bodies().tobytes() == b.bodies().tobytes(), i
    assert a.contacts().tobytes() == b.contacts().tobytes()
    # the plain step's callbacks are the split step's without PRE_SOLVE
    ea, eb = a.events(), b.events()
    assert np.array_equal(ea, eb[eb["type"] != native.EVENT_PRE_SOLVE])
    assert np.count_nonzero(eb["type"] == native.EVENT_PRE_SOLVE) > 10  # most bodies sleep by now


def test_presolve_split_state_errors():
    sc = scenes.mixed_container(60, 40.0, 80.0, seed=34)
    g = _gpu([sc])
    g.step(flags=FLAGS, n=30)
    s = np.zeros(1, native.SETTING_DTYPE)
    with pytest.raises(native.DbxError) as e:
        g.write_contact_settings(0, s)
    assert e.value.status == native.DBX_E_STATE
    with pytest.raises(native.DbxError) as e:
        g.step_finish()
    assert e.value.status == native.DBX_E_STATE
    g.step(flags=FLAGS | native.PRESOLVE)
    with pytest.raises(native.DbxError) as e:
        g.step(flags=FLAGS)
    assert e.value.status == native.DBX_E_STATE
    s["fixture_a"], s["fixture_b"] = 0, 1  # two static edges never touch
    with pytest.raises(native.DbxError) as e:
        g.write_contact_settings(0, s)
    assert e.value.status == native.DBX_E_INVALID_ARG
    g.step_finish()
    g.step(flags=FLAGS)
    g.sync()
