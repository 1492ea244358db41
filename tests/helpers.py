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


# ---------------------------------------------------------------------------
# bit-exact comparisons of GPU read-backs with the oracle's
# ---------------------------------------------------------------------------
BODY_FLOATS = ("px", "py", "qs", "qc", "cx", "cy", "a", "sleep_time", "vx", "vy", "w", "c0x", "c0y", "a0", "inv_mass")


def body_mismatch(bg: np.ndarray, bo: np.ndarray, types: np.ndarray | None = None) -> str | None:
    """None when every float field of every body has the same bits and the awake
    flags of the non-static bodies agree; otherwise a description of the first
    difference.  (dbox sets a static body's awake flag as a side effect of the islands
    it joins, b2island.d:241-279 and b2world.d:996; nothing in the step reads it.)"""
    if len(bg) != len(bo):
        return f"body count {len(bg)} != {len(bo)}"
    for f in BODY_FLOATS:
        a, b = bg[f].view(np.uint32), bo[f].view(np.uint32)
        bad = np.nonzero(a != b)[0]
        if bad.size:
            i = int(bad[0])
            d = np.abs(bg[f].astype(np.float64) - bo[f].astype(np.float64))
            return f"{bad.size} bodies differ in {f}, first {i}: {bg[f][i]!r} vs {bo[f][i]!r}, max |d| {d.max():.3e}"
    static = np.zeros(len(bg), bool) if types is None else np.asarray(types) == scenes.STATIC
    awake = ((bg["flags"] & 2) != (bo["flags"] & 2)) & ~static
    if awake.any():
        return f"{int(awake.sum())} bodies differ in the awake flag, first {int(np.nonzero(awake)[0][0])}"
    return None


def contact_mismatch(cg: np.ndarray, co: np.ndarray) -> str | None:
    """None when the contact lists are equal in order, fixtures, touching/enabled flags,
    point count, mixed material and, for contacts with points, manifold (type, normal,
    local points, feature ids) and the points' impulses, all bit for bit.  (A manifold
    without points keeps whatever type and normal its last evaluation left: the
    b2Collide* functions return before writing them, b2collidepolygon.d:135-150.)"""
    if len(cg) != len(co):
        return f"contact count {len(cg)} != {len(co)}"
    for f in ("fixture_a", "fixture_b", "point_count"):
        if not np.array_equal(cg[f], co[f]):
            return f"contacts differ in {f} at {int(np.nonzero(cg[f] != co[f])[0][0])}"
    if not np.array_equal(cg["flags"] & 6, co["flags"] & 6):
        return "contacts differ in touching/enabled"
    for f in ("friction", "restitution"):
        if cg[f].tobytes() != co[f].tobytes():
            return f"contacts differ in {f}"
    has = co["point_count"] > 0
    for f in ("type", "local_normal", "local_point"):
        if cg[f][has].tobytes() != co[f][has].tobytes():
            return f"contacts with points differ in {f}"
    for k in range(len(co)):
        n = int(co["point_count"][k])
        for f, w in (("point_local", 2 * n), ("normal_impulse", n), ("tangent_impulse", n), ("id", n)):
            if cg[f][k][:w].tobytes() != co[f][k][:w].tobytes():
                return f"contact {k} differs in {f}: {cg[f][k][:w]} vs {co[f][k][:w]}"
    return None


def varied_container(n: int, seed: int, width: float = 30.0, height: float = 40.0) -> scenes.Scene:
    """The mixed container of the batched workload with every per-body and per-fixture
    option the default scenes leave at its default: seeded collision categories,
    masks and groups (b2Filter), restitution U[0, 0.8], linear and angular damping
    U[0, 0.6], gravity scale U[0.5, 1.5], initial velocities, fixed rotation on some
    bodies and a few bullets."""
    s = scenes.mixed_container(n, width, height, seed)
    rng = scenes.Stream(seed ^ 0x5EED)
    b, f = s.bodies, s.fixtures
    dyn = b["type"] == scenes.DYNAMIC
    m = int(dyn.sum())
    b["linear_damping"][dyn] = rng.uniform(m, 0.0, 0.6)
    b["angular_damping"][dyn] = rng.uniform(m, 0.0, 0.6)
    b["gravity_scale"][dyn] = rng.uniform(m, 0.5, 1.5)
    b["vx"][dyn] = rng.uniform(m, -3.0, 3.0)
    b["vy"][dyn] = rng.uniform(m, -3.0, 3.0)
    b["w"][dyn] = rng.uniform(m, -2.0, 2.0)
    b["fixed_rotation"][dyn] = (rng.uniform(m) < 0.1).astype(np.int32)
    b["bullet"][dyn] = (rng.uniform(m) < 0.05).astype(np.int32)
    body_dyn = dyn[f["body"]]
    k = int(body_dyn.sum())
    f["restitution"][body_dyn] = rng.uniform(k, 0.0, 0.8)
    f["category"][body_dyn] = np.array([1, 2, 4], np.uint16)[rng.integers(k, 0, 2)]
    f["mask"][body_dyn] = np.array([0xFFFF, 0xFFFF, 0xFFFF ^ 2, 0xFFFF ^ 4, 1 | 4], np.uint16)[rng.integers(k, 0, 4)]
    f["group"][body_dyn] = np.array([0, 0, 0, 0, -1, -2, 1], np.int16)[rng.integers(k, 0, 6)]
    s.name = f"varied{n}"
    return s
