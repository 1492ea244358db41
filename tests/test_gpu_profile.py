"""dbx_profile: b2Profile's fields with b2World.Step's meaning (b2world.d:367-434,
930-1124; b2island.d:118-237)."""
import numpy as np
import pytest

from dbox_amd import native, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ordered", [True, False])
def test_profile_phases_sum_inside_solve(ordered):
    # solveInit + solveVelocity + solvePosition are phases of the island solve, which is
    # one part of b2World.Solve (solve); every phase is measured on a settling pyramid
    print(f"case: profile fields, pyramid, ordered={ordered}", flush=True)
    flags = native.WARM | native.SLEEP | native.CCD | (native.ORDERED if ordered else 0)
    h = native.StepWorlds()
    h.add_scene(scenes.pyramid())
    h.add_scene(scenes.mixed_container(200, 40.0, 80.0, seed=3))
    h.finalize()
    h.step(flags=flags, n=40)
    h.sync()
    h.profile_totals(reset=True)
    for _ in range(20):
        h.step(flags=flags)
        h.sync()
        p = h.profile()
        assert p["solveInit"] > 0 and p["solveVelocity"] > 0 and p["solvePosition"] > 0, p
        inner = p["solveInit"] + p["solveVelocity"] + p["solvePosition"]
        assert inner <= p["solve"] * 1.02 + 0.005, p
        assert p["broadphase"] <= p["solve"] and p["solve"] + p["collide"] <= p["step"] * 1.02 + 0.005, p
    tot, steps = h.profile_totals(reset=True)
    assert steps == 20
    assert tot["solvePosition"] > 0 and tot["solveVelocity"] > 0
    # several steps per sync: each step still gets its own phases
    h.step(flags=flags, n=5)
    h.sync()
    tot, steps = h.profile_totals(reset=True)
    assert steps == 5 and tot["solveVelocity"] > 0
    print(f"last step: {p}", flush=True)


def test_profile_phases_of_a_long_run_without_sync():
    # the island-solve phases of every step between two syncs are kept (up to
    # 4,096 steps): the same 600 steps timed without a sync and with a sync every
    # 50 steps give about the same phase totals (keeping only the last 256 steps'
    # phases would give the first less than half of the second)
    print("case: profile fields, 600 steps with and without syncs", flush=True)
    flags = native.WARM | native.SLEEP | native.CCD | native.ORDERED
    hs = []
    for _ in range(2):
        h = native.StepWorlds()
        h.add_scene(scenes.mixed_container(300, 40.0, 80.0, seed=5))
        h.finalize()
        h.step(flags=flags, n=20)
        h.sync()
        h.profile_totals(reset=True)
        hs.append(h)
    hs[0].step(flags=flags, n=600)
    hs[0].sync()
    for _ in range(12):
        hs[1].step(flags=flags, n=50)
        hs[1].sync()
    (ta, na), (tb, nb) = hs[0].profile_totals(reset=True), hs[1].profile_totals(reset=True)
    assert na == nb == 600
    inner = [t["solveInit"] + t["solveVelocity"] + t["solvePosition"] for t in (ta, tb)]
    print(f"phase totals without / with syncs: {inner[0]:.1f} / {inner[1]:.1f} ms", flush=True)
    assert inner[0] <= ta["solve"] * 1.02 + 0.005 * na, ta
    assert 0.75 * inner[1] < inner[0] < 1.33 * inner[1]
