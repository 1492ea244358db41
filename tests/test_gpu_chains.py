"""GPU parity of chain shapes (b2ChainShape, b2chainshape.d:24-263): the library
expands each chain into one edge proxy per segment (GetChildEdge), with the
chain's AABB rule (no radius), and names fixtures by proxy index.

* on the chain-terrain world (four chain pieces joined by ghost vertices, a loop
  chain around it, 60 bodies dropped on it, continuous physics on) every step's
  TOI event count and contacts (the same pairs in the same order with the same
  touching flags) equal the oracle's, the UpdatePairs pair set is bit-exact on
  steps without TOI events, fat AABBs agree, and bodies track the oracle;
* dbx_read_proxies maps proxy indices back to (fixture, child);
* batched chain worlds equal the same worlds stepped alone.
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


def test_chain_terrain_lockstep_with_oracle():
    print("case: chain terrain, 60 bodies, 90 steps", flush=True)
    s = scenes.chain_terrain(60)
    o = po.OracleWorld(s, continuous=True)
    g = _gpu([s])
    np.testing.assert_array_equal(g.fat_aabbs(), o.fat_aabbs())
    exact_steps = pair_steps = 0
    for i in range(90):
        o.step()
        g.step(flags=FLAGS)
        bo, bg = o.bodies(), g.bodies()
        dp = max(np.abs(bo["cx"] - bg["cx"]).max(), np.abs(bo["cy"] - bg["cy"]).max())
        if dp > 1e-4:
            print(f"states separate at step {i + 1} ({dp:.2e})")
            break
        # with continuous physics the oracle's last UpdatePairs is a TOI event's, when
        # the step has one; the GPU reports the discrete broadphase's pass
        assert int(g.toi_events()[0]) == o.toi_events(), i
        if o.toi_events() == 0:
            assert np.array_equal(g.pairs(), o.last_pairs()), i
            pair_steps += 1
        co, cg = o.contacts(), g.contacts()
        assert len(co) == len(cg), (i, len(co), len(cg))
        assert np.array_equal(cg["fixture_a"], co["fixture_a"]) and np.array_equal(cg["fixture_b"], co["fixture_b"])
        assert np.array_equal(cg["flags"] & 6, co["flags"] & 6)
        np.testing.assert_allclose(g.fat_aabbs(), o.fat_aabbs(), rtol=1e-6, atol=5e-5)
        exact_steps += 1
    print(f"contacts identical for {exact_steps} steps, pair sets for the {pair_steps} without TOI events")
    assert exact_steps >= 60 and pair_steps >= 40
    # some bodies rest on chain segments
    c = g.contacts()
    n_chain = scenes.proxy_count(s.fixtures[:5])
    assert np.count_nonzero((c["fixture_a"] < n_chain) & ((c["flags"] & 2) != 0)) > 10


def test_proxy_map_and_batched_chain_worlds():
    print("case: proxy map, 4 chain worlds batched vs single", flush=True)
    worlds = [scenes.chain_terrain(30, seed=200 + k) for k in range(4)]
    g = _gpu(worlds)
    fix, child = g.proxies(0)
    f0 = worlds[0].fixtures
    want_f, want_c = [], []
    for i, f in enumerate(f0):
        kids = scenes.proxy_count(f0[i:i + 1])
        want_f += [i] * kids
        want_c += list(range(kids))
    assert np.array_equal(fix, want_f) and np.array_equal(child, want_c)
    g.step(flags=FLAGS, n=60)
    for k in (0, 3):
        one = _gpu([worlds[k]])
        one.step(flags=FLAGS, n=60)
        assert g.bodies(k).tobytes() == one.bodies(0).tobytes(), k
        assert g.contacts(k).tobytes() == one.contacts(0).tobytes(), k
        one.close()
