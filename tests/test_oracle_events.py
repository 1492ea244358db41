"""The oracle's record of b2ContactListener callbacks (BeginContact / EndContact
from b2Contact.Update and b2ContactManager.Destroy, PostSolve from
b2Island.Report), checked against what fixes it without running dbox:

* PostSolve comes once per solved contact, in island order: the same pairs, in
  the same order, as the oracle's solve log (whose order the island tests pin);
* Begin and End alternate per contact, Begin first, and a Begin falls exactly on
  the step after which the contact reads as touching (End on the step after which
  it does not);
* a box resting on the ground: its contact's PostSolve normal impulses add up to
  m g dt once the stack has settled (the impulse that cancels one step of gravity).
"""
from __future__ import annotations

import numpy as np

from dbox_amd import scenes
from oracle import pyoracle as po

BEGIN, END, POST = 1, 2, 3


def _touching_pairs(o):
    c = o.contacts()
    t = (c["flags"] & 2) != 0
    return {(int(a), int(b)) for a, b in zip(c["fixture_a"][t], c["fixture_b"][t])}


def test_post_solve_follows_the_solve_order():
    sc = scenes.mixed_container(150, 40.0, 80.0, seed=21)
    body_of = sc.fixtures["body"].astype(np.int64)
    o = po.OracleWorld(sc)
    seen = 0
    for _ in range(90):
        o.step()
        e = o.events()
        post = e[e["type"] == POST]
        log = o.solve_log()  # (island, bodyA, bodyB) per solved contact, in solve order
        assert len(post) == len(log)
        if len(log):
            assert np.array_equal(body_of[post["fixture_a"]], log[:, 1])
            assert np.array_equal(body_of[post["fixture_b"]], log[:, 2])
            assert np.all((post["point_count"] >= 1) & (post["point_count"] <= 2))
            assert np.all(post["normal_impulse"][:, 0] >= 0.0)
            seen += len(log)
        # callbacks of the Collide phase come before any PostSolve
        if len(post):
            assert np.all(e["type"][: len(e) - len(post)] != POST)
    assert seen > 1000


def test_begin_and_end_alternate_and_match_touching():
    o = po.OracleWorld(scenes.mixed_container(150, 40.0, 80.0, seed=22))
    state = {}
    before = _touching_pairs(o)
    began = ended = 0
    for _ in range(150):
        o.step()
        e = o.events()
        after = _touching_pairs(o)
        for ev in e[e["type"] != POST]:
            key = (int(ev["fixture_a"]), int(ev["fixture_b"]))
            if ev["type"] == BEGIN:
                assert not state.get(key, False), key
                assert key not in before and key in after
                state[key] = True
                began += 1
            else:
                assert state.get(key, False), key
                assert key in before and key not in after
                state[key] = False
                ended += 1
        # every change of touching between two steps was reported
        assert {k for k, v in state.items() if v} == after
        before = after
    assert began > 100 and ended > 10


def test_resting_box_post_solve_is_one_step_of_gravity():
    sc = scenes.single_box_on_ground(y=0.5, hx=0.5, density=1.0)
    o = po.OracleWorld(sc, allow_sleep=False)
    for _ in range(120):
        o.step()
    post = o.events()
    post = post[post["type"] == POST]
    assert len(post) == 1
    m = 1.0 * (2 * 0.5) * (2 * 0.5)  # density x area
    total = float(post["normal_impulse"][0, : post["point_count"][0]].sum())
    assert abs(total - m * 10.0 / 60.0) < 1e-3 * m * 10.0 / 60.0
    assert np.all(np.abs(post["tangent_impulse"][0]) < 1e-6)


PRE = 4


def test_presolve_record_is_passive_and_covers_the_touching_contacts():
    # recording PreSolve calls changes nothing; with sleeping off every contact
    # that touches after Collide gets exactly one call, after its Begin
    sc = scenes.mixed_container(120, 40.0, 80.0, seed=31)
    quiet = po.OracleWorld(sc, allow_sleep=False)
    rec = po.OracleWorld(sc, allow_sleep=False)
    rec.set_presolve(1)
    calls = 0
    for _ in range(80):
        quiet.step()
        rec.step()
        assert rec.bodies().tobytes() == quiet.bodies().tobytes()
        e = rec.events()
        pre = e[e["type"] == PRE]
        keys = [(int(a), int(b)) for a, b in zip(pre["fixture_a"], pre["fixture_b"])]
        assert len(keys) == len(set(keys))
        assert set(keys) == _touching_pairs(rec)
        # a Begin comes before the PreSolve of the same contact
        order = {k: i for i, k in enumerate((int(a), int(b), int(t)) for a, b, t in
                                            zip(e["fixture_a"], e["fixture_b"], e["type"]))}
        for a, b in keys:
            if (a, b, BEGIN) in order:
                assert order[(a, b, BEGIN)] < order[(a, b, PRE)]
        assert np.all(pre["point_count"] <= 2)
        calls += len(keys)
    assert calls > 800


def test_presolve_policy_disables_and_changes_friction():
    # the test policy: (fixtureA + fixtureB) % 7 == 0 disables the contact for the
    # step (no PostSolve), == 1 sets friction 0.05: tangent impulses stay within
    # 0.1 x normal (the solver clamps against the previous iteration's normal
    # impulse, so a final ratio slightly above 0.05 occurs; the mixed default
    # frictions here are >= 0.2), and some contacts slide at that cone
    sc = scenes.mixed_container(150, 40.0, 80.0, seed=32)
    o = po.OracleWorld(sc)
    o.set_presolve(2)
    low_friction = disabled_seen = sliding = 0
    for _ in range(120):
        o.step()
        e = o.events()
        post = e[e["type"] == POST]
        k = (post["fixture_a"] + post["fixture_b"]) % 7
        assert not np.any(k == 0)
        pre = e[e["type"] == PRE]
        disabled_seen += int(np.count_nonzero((pre["fixture_a"] + pre["fixture_b"]) % 7 == 0))
        lf = post[k == 1]
        for ev in lf:
            n = int(ev["point_count"])
            t, nn = np.abs(ev["tangent_impulse"][:n]), ev["normal_impulse"][:n]
            assert np.all(t <= 0.1 * nn + 1e-6)
            sliding += int(np.count_nonzero((t > 0.04 * nn) & (nn > 1e-4)))
        low_friction += len(lf)
    assert disabled_seen > 50 and low_friction > 50 and sliding > 20
