"""Writes the golden fixtures of tests/golden from the CPU oracle.

The reference (dbox, D) ships no golden vectors and cannot be built here, so
these fixtures pin the oracle against later changes of itself: the CPU suite
checks that the oracle still reproduces them bit for bit, and the GPU suite
checks the device path against them.  Run from the repository root:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dbox_amd import scenes  # noqa: E402
from oracle import pyoracle as po  # noqa: E402
from tests.helpers import random_shape  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def pyramid_trace():
    """C1 (20-row pyramid, seed of config 1): body states at steps 1, 60, 600, the step's
    pair list at those steps, a digest of every step's pair list and the max penetration."""
    sc = scenes.pyramid(20)
    w = po.OracleWorld(sc)
    states, pairs, digests, pen = {}, {}, [], []
    for step in range(1, 601):
        w.step()
        p = w.last_pairs()
        digests.append(hashlib.sha256(p.tobytes()).hexdigest()[:16])
        pen.append(w.max_penetration())
        if step in (1, 60, 600):
            states[step] = w.bodies()
            pairs[step] = p
    np.savez_compressed(os.path.join(HERE, "pyramid20.npz"),
                        bodies_1=states[1], bodies_60=states[60], bodies_600=states[600],
                        pairs_1=pairs[1], pairs_60=pairs[60], pairs_600=pairs[600],
                        pair_digests=np.array(digests), max_penetration=np.array(pen, np.float32))


def manifolds(n=2000, seed=20240):
    """n random shape pairs at random transforms and their oracle manifolds."""
    rng = np.random.default_rng(seed)
    fa, fb, xa, xb, out = [], [], [], [], []
    while len(out) < n:
        a, b = random_shape(rng), random_shape(rng)
        ta = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-3.2, 3.2)], np.float32)
        tb = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-3.2, 3.2)], np.float32)
        m = po.collide(a, ta, b, tb)
        if m is None:
            continue
        fa.append(a)
        fb.append(b)
        xa.append(ta)
        xb.append(tb)
        out.append(m)
    np.savez_compressed(os.path.join(HERE, "manifolds.npz"), fixture_a=np.array(fa), fixture_b=np.array(fb),
                        xf_a=np.array(xa), xf_b=np.array(xb), manifold=np.array(out))


if __name__ == "__main__":
    pyramid_trace()
    manifolds()
    print("written to", HERE)
