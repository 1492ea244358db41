"""The kernels' arithmetic is the reference's, bit for bit.

The oracle runs with b2Rot.Set evaluated as the device library evaluates it
(sin and cos in double, rounded once: ``exact_rot=True``); every other operation
of both is the reference's float operation order.  With that one libm call made
equal, the GPU path must give the oracle's bodies (every float field), contact
lists (order, flags, manifolds, feature ids, impulses), broadphase pair sets and
TOI event counts bit for bit, at every compared step:

  * config 1, the 20-row pyramid, 600 steps with dbox's defaults (continuous physics on);
  * config 2, 20,000 bodies in the 200 m x 400 m container, 60 steps (device-wide
    broadphase, Collide and first TOI search), and the same world with every 40th
    body thrown down at 120 m/s, so that the device-wide continuous pass has events;
  * config 4, the 50,000-box pyramid, 30 steps with continuous physics on;
  * config 3, all 4096 worlds in one batch for 150 steps, against the oracle on the
    five largest worlds (the lead tier) and one world of each other tier, and against
    the same eight worlds stepped alone;
  * a container exercising contact filtering, restitution, damping, gravity scale,
    fixed rotation and bullets under a varying time step.

The reference-faithful comparison (glibc sinf/cosf in the oracle) is in
test_gpu_parity.py; there trajectories differ by the libm's last-bit rounding
amplified by the scene.
"""
from __future__ import annotations

import numpy as np
import pytest

from dbox_amd import native, scenes
from oracle import pyoracle as po
from tests.helpers import body_mismatch, contact_mismatch, varied_container

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0
DEFAULTS = native.WARM | native.SLEEP | native.CCD | native.ORDERED


def _gpu(scene_list):
    h = native.StepWorlds()
    for s in scene_list:
        h.add_scene(s)
    h.finalize()
    return h


def _check(step, g, o, world=0, contacts=True, pairs=True):
    m = body_mismatch(g.bodies(world), o.bodies(), o.bodies_def["type"])
    assert m is None, f"step {step}: {m}"
    if contacts:
        m = contact_mismatch(g.contacts(world), o.contacts())
        assert m is None, f"step {step}: {m}"
    if pairs:
        assert np.array_equal(g.pairs(world), o.last_pairs()), f"step {step}: pairs"


def _lockstep(scene, steps, every, flags=DEFAULTS, dts=None, label=""):
    o = po.OracleWorld(scene, continuous=bool(flags & native.CCD), exact_rot=True)
    g = _gpu([scene])
    events = 0
    for i in range(steps):
        dt = DT if dts is None else float(dts[i % len(dts)])
        o.step(dt)
        g.step(dt, flags=flags)
        gi = int(g.toi_events()[0]) if flags & native.CCD else 0
        assert gi == o.toi_events(), f"step {i + 1}: TOI events {gi} vs {o.toi_events()}"
        events += gi
        if (i + 1) % every == 0 or i + 1 == steps:
            # the pair set is the oracle's last UpdatePairs: a TOI event's when the pass had one
            _check(i + 1, g, o, pairs=o.toi_events() == 0)
    print(f"{label}: {steps} steps bit-identical, {len(o.contacts())} contacts, {events} TOI events", flush=True)
    return o, g


def test_pyramid_600_steps_bit_identical():
    print("case: config 1 pyramid, 600 steps, dbox defaults, exact sin/cos", flush=True)
    o, g = _lockstep(scenes.pyramid(), 600, 10, label="config 1")
    assert len(o.contacts()) > 500


def test_big_mixed_world_60_steps_bit_identical():
    print("case: config 2, 20,000 bodies, 60 steps, exact sin/cos", flush=True)
    _lockstep(scenes.big_mixed_world(), 60, 10, label="config 2")


def test_big_mixed_world_with_projectiles_continuous_pass():
    # every 40th body thrown down at 120 m/s: the device-wide first TOI search and its
    # events on a world too big for one workgroup
    print("case: config 2 with projectiles, 24 steps, continuous pass, exact sin/cos", flush=True)
    s = scenes.big_mixed_world()
    fast = np.arange(1, len(s.bodies), 40)
    s.bodies["vy"][fast] = -120.0
    o, g = _lockstep(s, 24, 4, label="config 2 + projectiles")


def test_big_pyramid_30_steps_bit_identical():
    print("case: config 4, 50,000 boxes, 30 steps, exact sin/cos", flush=True)
    _lockstep(scenes.big_pyramid(), 30, 6, label="config 4")


def test_varied_container_variable_dt_bit_identical():
    print("case: filters, restitution, damping, bullets, varying dt, 240 steps", flush=True)
    dts = np.array([1 / 60, 1 / 60, 1 / 30, 1 / 120, 1 / 45, 1 / 60, 1 / 90], np.float32)
    for seed in (3, 4):
        o, g = _lockstep(varied_container(250, seed), 240, 8, dts=dts, label=f"varied seed {seed}")
        # the scene exercised what it is for
        c = o.contacts()
        assert (c["restitution"] > 0).any()


def test_config3_all_4096_worlds_sampled_against_oracle_and_single():
    print("case: config 3, 4096 worlds, 150 steps, exact sin/cos", flush=True)
    worlds = scenes.batched_worlds(4096)
    g = _gpu(worlds)
    tiers, order = g.tiers()
    lead = [int(order[i]) for i in range(int(tiers[0, 1]))]
    sample = lead + [int(order[int(t[0]) + int(t[1]) // 2]) for t in tiers[1:]]
    print(f"tiers {tiers.tolist()}, sampled worlds {sample}", flush=True)
    assert len(tiers) >= 4 and len(sample) >= 8
    ref = {w: po.OracleWorld(worlds[w], continuous=True, exact_rot=True) for w in sample}
    for i in range(150):
        g.step(flags=DEFAULTS)
        for w in sample:
            ref[w].step()
        if (i + 1) % 30 == 0:
            ev = g.toi_events()
            for w in sample:
                assert int(ev[w]) == ref[w].toi_events(), (i + 1, w)
                _check(i + 1, g, ref[w], world=w, pairs=False)
    for w in sample:
        single = _gpu([worlds[w]])
        single.step(flags=DEFAULTS, n=150)
        assert g.bodies(w).tobytes() == single.bodies(0).tobytes(), w
        assert g.contacts(w).tobytes() == single.contacts(0).tobytes(), w
        single.close()
    touching = sum(int(((ref[w].contacts()["flags"] & 2) != 0).sum()) for w in sample)
    print(f"config 3: {len(sample)} worlds bit-identical to the oracle and to single runs; {touching} touching contacts")
    assert touching > 1000


@pytest.mark.parametrize("which,steps", [("config 1 pyramid", 60), ("config 2 container", 40), ("config 4 pyramid", 24)])
def test_residual_mode_at_max_iterations_equals_the_plain_step(which, steps):
    # RESIDUAL mode (no counterpart in dbox) sweeps until the velocity residual
    # falls to the tolerance; with a negative tolerance and a cap of 8 it runs all
    # eight sweeps, which make the same updates as the plain step's 8.  That pins
    # its loops on the three solve paths: the one-wave LDS solve (config 1), the
    # global-memory sweeps (config 2) and the compact LDS sweeps of a big world
    # (config 4).
    print(f"case: RESIDUAL at 8 iterations vs plain, {which}, {steps} steps", flush=True)
    scene = {"config 1 pyramid": scenes.pyramid, "config 2 container": scenes.big_mixed_world,
             "config 4 pyramid": scenes.big_pyramid}[which]()
    plain, resid = _gpu([scene]), _gpu([scene])
    resid.set_residual(-1.0, 8)
    solved = 0
    for i in range(steps):
        plain.step(DT, 8, 3, DEFAULTS)
        resid.step(DT, 8, 3, DEFAULTS | native.RESIDUAL)
        it, _ = resid.solver_stats()
        assert int(it[0]) in (0, 8), f"step {i + 1}: {int(it[0])} sweeps"
        solved += int(it[0] == 8)
        assert resid.bodies(0).tobytes() == plain.bodies(0).tobytes(), f"step {i + 1}: bodies"
    assert solved > 0
    assert resid.contacts(0).tobytes() == plain.contacts(0).tobytes()


@pytest.mark.parametrize("which,steps", [("config 2 container", 40), ("config 4 pyramid", 30)])
def test_island_groups_equal_one_workgroup(which, steps, monkeypatch):
    # A big world's islands are solved by many workgroups at once (each a range of
    # islands with about 1,024 contact slots' worth of constraints); with
    # DBX_SOLVE_UNGROUPED set a step solves them all on one workgroup.  Islands are
    # independent in b2World.Solve, so every step must give the same bodies and
    # contacts bit for bit, including steps where trailing groups hold no island.
    print(f"case: island groups vs one workgroup, {which}, {steps} steps", flush=True)
    scene = {"config 2 container": scenes.big_mixed_world, "config 4 pyramid": scenes.big_pyramid}[which]()
    grouped, single = _gpu([scene]), _gpu([scene])
    for i in range(steps):
        monkeypatch.delenv("DBX_SOLVE_UNGROUPED", raising=False)
        grouped.step(DT, 8, 3, DEFAULTS)
        monkeypatch.setenv("DBX_SOLVE_UNGROUPED", "1")
        single.step(DT, 8, 3, DEFAULTS)
        assert int(grouped.toi_events()[0]) == int(single.toi_events()[0]), f"step {i + 1}: TOI events"
        assert grouped.bodies(0).tobytes() == single.bodies(0).tobytes(), f"step {i + 1}: bodies"
    assert grouped.contacts(0).tobytes() == single.contacts(0).tobytes()


@pytest.mark.parametrize("vel,warm", [(8, True), (3, False), (1, True)])
def test_overlapped_sweeps_equal_sequential_sweeps(vel, warm, monkeypatch):
    # The lead tier's workgroups may run the warm start and the velocity
    # iterations overlapped on two waves (sweep s + 1 trailing sweep s only as far
    # as each body's updates require), which they do where the schedule makes it
    # pay; DBX_SWEEPS_OVERLAP_ALWAYS overlaps them at every step and
    # DBX_SWEEPS_SEQUENTIAL runs them one after the other on one wave.  The
    # overlapped order keeps every two updates of a body in the sequential order,
    # so bodies and contacts must agree bit for bit across the three, with and
    # without warm starting and for any number of iterations.
    print(f"case: overlapped vs sequential sweeps, {vel} iterations, warm {warm}", flush=True)
    worlds = scenes.batched_worlds(4096)
    sizes = np.array([len(s.bodies) for s in worlds])
    pick = list(np.argsort(-sizes, kind="stable")[:6]) + list(range(0, 4096, 128))
    chosen = [worlds[int(i)] for i in pick]
    flags = DEFAULTS if warm else DEFAULTS & ~native.WARM
    modes = {"default": {}, "always": {"DBX_SWEEPS_OVERLAP_ALWAYS": "1"}, "sequential": {"DBX_SWEEPS_SEQUENTIAL": "1"}}
    hs = {m: _gpu(chosen) for m in modes}
    for i in range(160):
        for m, env in modes.items():
            for var in ("DBX_SWEEPS_OVERLAP_ALWAYS", "DBX_SWEEPS_SEQUENTIAL"):
                monkeypatch.delenv(var, raising=False)
            for var, val in env.items():
                monkeypatch.setenv(var, val)
            hs[m].step(DT, vel, 3, flags)
        if i % 20 == 19:
            for k in range(len(chosen)):
                ref = hs["sequential"].bodies(k).tobytes()
                for m in ("default", "always"):
                    assert hs[m].bodies(k).tobytes() == ref, f"{m}, step {i + 1}, world {k}: bodies"
    for k in range(len(chosen)):
        ref = hs["sequential"].contacts(k).tobytes()
        for m in ("default", "always"):
            assert hs[m].contacts(k).tobytes() == ref, f"{m}, world {k}: contacts"
    it, _ = hs["always"].solver_stats()
    assert int(it[0]) == vel
