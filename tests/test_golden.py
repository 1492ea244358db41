"""Golden fixtures (tests/golden, written by tests/golden/make_golden.py from the
oracle): the CPU oracle must keep reproducing them bit for bit, and the GPU path
is checked against them at the parity tiers."""
from __future__ import annotations

import hashlib
import os

import numpy as np
import pytest

from dbox_amd import scenes
from oracle import pyoracle as po

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def test_oracle_reproduces_pyramid_trace():
    g = _load("pyramid20.npz")
    w = po.OracleWorld(scenes.pyramid(20))
    for step in range(1, 61):
        w.step()
        p = w.last_pairs()
        assert hashlib.sha256(p.tobytes()).hexdigest()[:16] == g["pair_digests"][step - 1], step
        assert np.float32(w.max_penetration()) == g["max_penetration"][step - 1], step
        if step in (1, 60):
            assert w.bodies().tobytes() == g[f"bodies_{step}"].tobytes(), step
            assert np.array_equal(p, g[f"pairs_{step}"])


def test_oracle_reproduces_manifolds():
    g = _load("manifolds.npz")
    for i in range(0, len(g["manifold"]), 7):
        m = po.collide(g["fixture_a"][i], g["xf_a"][i], g["fixture_b"][i], g["xf_b"][i])
        assert m.tobytes() == g["manifold"][i].tobytes(), i


@pytest.mark.gpu
def test_device_matches_golden_pyramid():
    from dbox_amd import native
    print("case: device pyramid vs golden states (steps 1, 60) and pairs", flush=True)
    g = _load("pyramid20.npz")
    h = native.StepWorlds()
    h.add_scene(scenes.pyramid(20))
    h.finalize()
    flags = native.WARM | native.SLEEP | native.ORDERED
    h.step(flags=flags)
    assert np.array_equal(h.pairs(0), g["pairs_1"])
    h.step(flags=flags, n=59)
    b, ref = h.bodies(0), g["bodies_60"]
    assert np.abs(b["cx"] - ref["cx"]).max() < 1e-3 and np.abs(b["cy"] - ref["cy"]).max() < 1e-3
    assert np.abs(b["a"] - ref["a"]).max() < 1e-3


@pytest.mark.gpu
def test_device_matches_golden_manifolds():
    from dbox_amd import native
    print("case: device narrowphase vs 2000 golden manifolds", flush=True)
    g = _load("manifolds.npz")
    h = native.StepWorlds()
    out = native.collide_pairs(h, g["fixture_a"], g["xf_a"], g["fixture_b"], g["xf_b"])
    m = g["manifold"]
    # out columns: type, pointCount, swapped, 0, localNormal, localPoint, points' local (4), ids (2),
    # world normal (14:16), world points (16:20), separations (20:22)
    cref, cdev = m["point_count"], out[:, 1].astype(np.int32)
    agree = cref == cdev
    # a near-tangent pair may flip between 0 and 1 points on a 1-ulp sin/cos difference
    for i in np.nonzero(~agree)[0]:
        assert cref[i] == 0 or cdev[i] == 0, i
        seps = m["separations"][i, :cref[i]] if cref[i] else out[i, 20:20 + cdev[i]]
        assert np.abs(seps).min() < 0.02, i
    assert agree.mean() > 0.99
    for j in range(2):
        has = agree & (cref > j)
        assert np.abs(out[has, 16 + 2 * j:18 + 2 * j] - m["points"][has, 2 * j:2 * j + 2]).max() < 1e-5, j
        assert np.abs(out[has, 20 + j] - m["separations"][has, j]).max() < 1e-5, j
    k = agree & (cref > 0)
    assert np.abs(out[k, 14:16] - m["normal"][k]).max() < 1e-5
