"""GPU checks at the full sizes of the large configurations: the 20,000-body mixed
world (config 2), the 50,000-box pyramid (config 4) and the 16,384 stack worlds
(config 5).  The oracle runs the single big worlds for their first steps, which
it finishes in well under a second, and every check there is exact or at the
manifold tier; longer runs are checked through properties that do not depend on
size (batched == single bit for bit, finiteness, bounded penetration).
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


@pytest.mark.parametrize("config", ["C2 mixed 20k", "C4 pyramid 50k"])
def test_big_world_first_steps_match_oracle(config):
    # per step: UpdatePairs pair set bit-exact, the same contacts in list order with
    # the same touching/enabled flags, fat AABBs and bodies within 5e-5 m or 1e-6
    # relative (the worlds span 400 m, where one float ulp is 3e-5 m, and the sin/cos
    # rounding differences of 16 steps of contact reach 1.3e-5 m)
    print(f"case: {config}, lockstep", flush=True)
    s = scenes.big_mixed_world() if config.startswith("C2") else scenes.big_pyramid()
    o = po.OracleWorld(s)
    g = _gpu([s])
    for i in range(16):  # the pyramid's rows first touch at step 13
        o.step()
        g.step(flags=FLAGS)
        assert np.array_equal(g.pairs(), o.last_pairs()), i
        co, cg = o.contacts(), g.contacts()
        assert len(co) == len(cg), (i, len(co), len(cg))
        assert np.array_equal(cg["fixture_a"], co["fixture_a"]) and np.array_equal(cg["fixture_b"], co["fixture_b"])
        assert np.array_equal(cg["flags"] & 6, co["flags"] & 6)
        np.testing.assert_allclose(g.fat_aabbs(), o.fat_aabbs(), rtol=1e-6, atol=5e-5)
        bo, bg = o.bodies(), g.bodies()
        for f in ("cx", "cy", "vx", "vy"):
            np.testing.assert_allclose(bg[f], bo[f], rtol=1e-6, atol=5e-5)
    print(f"{config}: {len(co)} contacts, {int((co['flags'] & 2).astype(bool).sum())} touching at step 16")


def test_stack_worlds_full_size_batched_equals_single():
    # config 5 at its full size (16,384 worlds, 1.65 M bodies): three worlds stepped
    # alone give the same bits as inside the batch, and the batch stays physical
    print("case: C5 16384 stack worlds, 60 steps", flush=True)
    worlds = scenes.stack_worlds(16384)
    g = _gpu(worlds)
    g.step(flags=FLAGS, n=60)
    it, pen = g.solver_stats()
    assert np.all(np.isfinite(pen))
    # tall stacks sink under 8 velocity iterations; the deepest world must not sink
    # more than the oracle's run of the same world by more than b2_linearSlop
    worst = int(np.argmax(pen))
    o = po.OracleWorld(worlds[worst])
    o.step(n=60)
    print(f"C5: max penetration over the batch {float(pen.max()):.4f} m (world {worst}); "
          f"oracle on that world {o.max_penetration():.4f} m")
    assert float(pen[worst]) <= o.max_penetration() + 0.005
    for k in (0, 7777, 16383):
        single = _gpu([worlds[k]])
        single.step(flags=FLAGS, n=60)
        assert g.bodies(k).tobytes() == single.bodies(0).tobytes(), k
        assert g.contacts(k).tobytes() == single.contacts(0).tobytes(), k
        single.close()
