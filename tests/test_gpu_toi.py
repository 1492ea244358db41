"""GPU parity of the continuous pass (b2World.SolveTOI, b2world.d:1127-1452;
b2TimeOfImpact, b2timeofimpact.d:67-295; b2Island.SolveTOI, b2island.d:282-416)
against the oracle with continuous physics on.

* the fast box over a thin edge: per step the same number of TOI events and
  bodies within 1e-5 m of the oracle;
* projectiles in a closed room (dozens of events per step): per step the same
  event counts while the two runs agree to 1e-3 m (the trajectories separate by
  sin/cos rounding, and a TOI root moves with them), and no body leaves the room
  on either side, where without the pass most of them do;
* the C1 pyramid with the pass on tracks the oracle like the discrete step;
* batched worlds equal the same worlds stepped alone, and a replayed hipGraph
  equals the eager step, bit for bit;
* a TOI event whose moved proxy's query finds a few hundred proxies tracks the
  oracle, and one whose query finds more than the pass's per-event pair list
  holds (kToiPairCap) is reported as a capacity error naming that list.
"""
from __future__ import annotations

import numpy as np
import pytest

from dbox_amd import native, scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
FLAGS = native.WARM | native.SLEEP | native.ORDERED | native.CCD


def _gpu(scene_list, **kw):
    h = native.StepWorlds(**kw)
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


def _outside(b, size):
    h = 0.5 * size
    return (np.abs(b["cx"]) > h + 0.5) | (b["cy"] < -0.5) | (b["cy"] > size + 0.5)


def test_fast_box_stops_on_the_edge_like_the_oracle():
    print("case: fast box over an edge, continuous", flush=True)
    s = _fast_box()
    o = po.OracleWorld(s, continuous=True)
    g = _gpu([s])
    for i in range(30):
        o.step()
        g.step(flags=FLAGS)
        assert int(g.toi_events()[0]) == o.toi_events(), i
        bo, bg = o.bodies(), g.bodies()
        for f in ("cx", "cy", "vx", "vy"):
            np.testing.assert_allclose(bg[f], bo[f], rtol=0, atol=1e-5, err_msg=f"step {i + 1} {f}")
    assert 0.25 <= float(g.bodies()["cy"][1]) <= 0.3


def test_projectiles_stay_in_the_room_and_track_the_oracle():
    print("case: projectile room, 40 bodies, 60 steps, continuous", flush=True)
    s = scenes.projectile_room(40)
    size = s.meta["size"]
    o = po.OracleWorld(s, continuous=True)
    g = _gpu([s])
    compared = total = 0
    tracking = True
    for i in range(60):
        o.step()
        g.step(flags=FLAGS)
        eo, eg = o.toi_events(), int(g.toi_events()[0])
        total += eg
        bo, bg = o.bodies(), g.bodies()
        if tracking:
            assert eg == eo, (i, eg, eo)
            compared += 1
            dp = np.hypot(bg["cx"] - bo["cx"], bg["cy"] - bo["cy"]).max()
            if dp > 1e-3:
                print(f"runs separate after step {i + 1} (max |dx| {dp:.2e})")
                tracking = False
        assert not _outside(bg[1:], size).any(), i
        assert not _outside(bo[1:], size).any(), i
    print(f"compared event counts over {compared} steps; {total} GPU events in 60 steps")
    assert compared >= 3 and total > 50
    # the same worlds without the pass lose most of their bodies
    g2 = _gpu([s])
    g2.step(flags=FLAGS & ~native.CCD, n=60)
    assert _outside(g2.bodies()[1:], size).sum() > 20


def test_pyramid_with_continuous_pass_tracks_the_oracle():
    print("case: C1 pyramid, 120 steps, continuous", flush=True)
    s = scenes.pyramid()
    o = po.OracleWorld(s, continuous=True)
    g = _gpu([s])
    for i in range(120):
        o.step()
        g.step(flags=FLAGS)
        assert int(g.toi_events()[0]) == o.toi_events(), i
    bo, bg = o.bodies(), g.bodies()
    dp = np.hypot(bg["cx"] - bo["cx"], bg["cy"] - bo["cy"])
    print(f"pyramid continuous: max |dx| {dp.max():.2e}")
    assert dp.max() < 1e-3


def test_continuous_batched_equals_single_and_graph_equals_eager():
    print("case: continuous, 6 projectile worlds batched vs single, graph vs eager", flush=True)
    worlds = [scenes.projectile_room(24, seed=100 + k) for k in range(6)]
    batch = _gpu(worlds)
    batch.step(flags=FLAGS, n=40)
    for k in (0, 3, 5):
        one = _gpu([worlds[k]])
        one.step(flags=FLAGS, n=40)
        assert batch.bodies(k).tobytes() == one.bodies(0).tobytes(), k
        assert batch.contacts(k).tobytes() == one.contacts(0).tobytes(), k
        one.close()
    gr = _gpu(worlds)
    gr.set_graph(True)
    gr.step(flags=FLAGS, n=40)
    for k in range(6):
        assert gr.bodies(k).tobytes() == batch.bodies(k).tobytes(), k



def _box_bouncing_over_pins(pitch):
    """A 6 m x 6 m box (no gravity, restitution 1) thrown at 120 m/s into a
    vertical wall edge 0.6 m ahead; behind it tiny static sensor pins on a square
    grid of the given pitch.  The TOI event bounces the box back beyond where the
    step started, out of the fat AABB of its discrete sweep, so the event moves
    its proxy and queries the pins its new fat AABB covers.  The pins' own fat
    AABBs (0.2 m wide) never overlap one another, and sensors take no part in the
    continuous pass."""
    xs = np.arange(-12.0, -0.2, pitch)
    ys = np.arange(-3.2, 3.2, pitch)
    n_pins = len(xs) * len(ys)
    b = scenes.body_defs(2 + n_pins)
    f = scenes.fixture_defs(2 + n_pins)
    b[0]["type"] = scenes.STATIC
    f[0]["body"] = 0
    f[0]["shape"] = scenes.SHAPE_EDGE
    f[0]["p"][:4] = (0.0, -10.0, 0.0, 10.0)
    b[1]["type"] = scenes.DYNAMIC
    b[1]["px"] = -3.6
    b[1]["vx"] = 120.0
    b[1]["gravity_scale"] = 0.0
    f[1]["body"] = 1
    f[1]["shape"] = scenes.SHAPE_BOX
    f[1]["p"][:2] = (3.0, 3.0)
    f[1]["density"] = 1.0
    f[1]["restitution"] = 1.0
    gx, gy = np.meshgrid(xs, ys)
    b[2:]["type"] = scenes.STATIC
    b[2:]["px"] = gx.ravel()
    b[2:]["py"] = gy.ravel()
    f[2:]["body"] = np.arange(2, 2 + n_pins)
    f[2:]["shape"] = scenes.SHAPE_CIRCLE
    f[2:]["radius"] = 0.001
    f[2:]["flags"] = scenes.SENSOR
    return scenes.Scene(b, f, name=f"box_bouncing_over_pins_{pitch}")


def test_event_query_below_the_list_cap_tracks_the_oracle():
    # about 350 proxies in the event's query: under the pass's per-event pair
    # list (512), so the GPU keeps the reference's AddPair order
    print("case: box bouncing over sensor pins at 0.42 m, continuous", flush=True)
    s = _box_bouncing_over_pins(0.42)
    o = po.OracleWorld(s, continuous=True)
    g = _gpu([s])
    events = 0
    for i in range(10):
        o.step()
        g.step(flags=FLAGS)
        assert int(g.toi_events()[0]) == o.toi_events(), i
        events += o.toi_events()
        bo, bg = o.bodies(), g.bodies()
        for fld in ("cx", "cy", "vx", "vy"):
            np.testing.assert_allclose(bg[fld], bo[fld], rtol=0, atol=1e-5, err_msg=f"step {i + 1} {fld}")
        if i == 0:
            # the event's query made contacts: more than the discrete step alone has
            d = po.OracleWorld(s, continuous=False)
            d.step()
            assert bg["vx"][1] < 0.0
            assert len(g.contacts()) == len(o.contacts()) > len(d.contacts())
    assert events >= 1
    g.sync()
    assert len(g.contacts()) == len(o.contacts())


def test_event_query_over_the_list_cap_is_reported():
    # over a thousand proxies in the event's query: more pairs than the per-event
    # list holds; the pass drops them and the step reports a capacity error
    # naming that list
    print("case: box bouncing over sensor pins at 0.21 m, continuous, pair list cap", flush=True)
    g = _gpu([_box_bouncing_over_pins(0.21)])
    g.step(flags=FLAGS)
    with pytest.raises(native.DbxError) as err:
        g.sync()
    assert err.value.status == native.DBX_E_CAPACITY
    assert "continuous-pass" in str(err.value)
