"""Shared helpers for the tests: single fixture definitions and comparisons."""
from __future__ import annotations

import numpy as np

from dbox_amd import scenes


def box(hx, hy, density=1.0, friction=0.6, center=None, angle=0.0):
    f = scenes.fixture_defs(1)[0]
    f["shape"] = scenes.SHAPE_BOX
    f["p"][:2] = (hx, hy)
    if center is not None:
        f["p"][2:5] = (center[0], center[1], angle)
        f["flags"] = 8
    f["density"] = density
    f["friction"] = friction
    return f


def circle(r, center=(0.0, 0.0), density=1.0, friction=0.6):
    f = scenes.fixture_defs(1)[0]
    f["shape"] = scenes.SHAPE_CIRCLE
    f["radius"] = r
    f["p"][:2] = center
    f["density"] = density
    f["friction"] = friction
    return f


def edge(x1, y1, x2, y2, v0=None, v3=None):
    f = scenes.fixture_defs(1)[0]
    f["shape"] = scenes.SHAPE_EDGE
    f["p"][:4] = (x1, y1, x2, y2)
    flags = 0
    if v0 is not None:
        f["p"][4:6] = v0
        flags |= 2
    if v3 is not None:
        f["p"][6:8] = v3
        flags |= 4
    f["flags"] = flags
    return f


def polygon(points, density=1.0, friction=0.6):
    f = scenes.fixture_defs(1)[0]
    f["shape"] = scenes.SHAPE_POLYGON
    pts = np.asarray(points, np.float32).reshape(-1, 2)
    f["count"] = len(pts)
    f["p"][: 2 * len(pts)] = pts.reshape(-1)
    f["density"] = density
    f["friction"] = friction
    return f


def random_shape(rng: np.random.Generator):
    """A random fixture of any supported kind (as the scene generators make them)."""
    k = rng.integers(0, 4)
    if k == 0:
        return circle(float(rng.uniform(0.2, 0.8)))
    if k == 1:
        return box(float(rng.uniform(0.2, 1.0)), float(rng.uniform(0.2, 1.0)))
    if k == 2:
        n = int(rng.integers(3, 9))
        r = rng.uniform(0.25, 1.0)
        ang = 2 * np.pi * (np.arange(n) + rng.uniform(-0.3, 0.3, n)) / n
        return polygon(np.stack([r * np.cos(ang), r * np.sin(ang)], 1))
    L = rng.uniform(1.0, 4.0)
    a = rng.uniform(0, 2 * np.pi)
    c, s = np.cos(a) * L / 2, np.sin(a) * L / 2
    v0 = (float(-c * 2.5), float(-s * 2.5)) if rng.random() < 0.5 else None
    v3 = (float(c * 2.5 + rng.uniform(-0.5, 0.5)), float(s * 2.5)) if rng.random() < 0.5 else None
    return edge(float(-c), float(-s), float(c), float(s), v0, v3)


def scene_of(bodies: list[tuple], fixtures: list) -> scenes.Scene:
    """bodies: (type, x, y, angle[, vx, vy, w]); fixtures: fixture records with 'body' set."""
    b = scenes.body_defs(len(bodies))
    for i, t in enumerate(bodies):
        b[i]["type"], b[i]["px"], b[i]["py"], b[i]["angle"] = t[:4]
        if len(t) > 4:
            b[i]["vx"], b[i]["vy"], b[i]["w"] = t[4:7]
    f = np.array(fixtures, scenes.FIXTURE_DTYPE)
    return scenes.Scene(b, f)
