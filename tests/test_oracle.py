"""Pins of the CPU oracle against closed forms, brute force and the one
known-answer scene of the reference (examples/hello_world/hello_world.d).

The reference ships no tests or golden vectors (no unittest blocks, no test
data), so every check here is derived from the mathematics or from the
documented behaviour of the example programs.
"""
import math

import numpy as np
import pytest

from dbox_amd import scenes
from oracle import pyoracle as po
from tests.helpers import box, circle, edge, polygon, random_shape, scene_of

DT = 1.0 / 60.0


def test_hello_world_rests_at_expected_height():
    # hello_world.d:31-102: a 2x2 box dropped from y=4 onto a ground box whose top is y=0,
    # 60 steps at 6/2 iterations.  At rest the centre sits at 1.0 m plus the contact slop.
    w = po.OracleWorld(scenes.hello_world())
    for _ in range(60):
        w.step(DT, 6, 2)
    b = w.bodies()[1]
    assert abs(b["px"]) < 0.01
    assert abs(b["py"] - 1.01) < 0.01
    assert abs(b["a"]) < 0.01


def test_free_fall_matches_semi_implicit_euler_bit_exactly():
    # b2island.d:82-116 and 168-200 with no contact: v += h*g; c += h*v, in float32.
    w = po.OracleWorld(scenes.hello_world())
    h = np.float32(DT)
    v = np.float32(0.0)
    y = np.float32(4.0)
    for _ in range(30):  # the box is still 2 m above the ground after 30 steps
        w.step(DT, 6, 2)
        v = np.float32(v + np.float32(h * np.float32(np.float32(1.0) * np.float32(-10.0) + np.float32(0.25) * np.float32(0.0))))
        v = np.float32(v * np.float32(np.float32(1.0) / np.float32(1.0 + h * np.float32(0.0))))
        y = np.float32(y + np.float32(h * v))
        b = w.bodies()[1]
        assert b["py"] == y and b["vy"] == v


@pytest.mark.parametrize("hx,hy,rho", [(0.5, 0.5, 5.0), (1.0, 0.25, 2.0), (0.3, 0.7, 1.3)])
def test_box_mass_and_inertia_closed_form(hx, hy, rho):
    _, m = po.shape_geometry(box(hx, hy, density=rho))
    mass = rho * 4 * hx * hy
    assert m[0] == pytest.approx(mass, rel=1e-6)
    assert abs(m[1]) < 1e-6 and abs(m[2]) < 1e-6
    assert m[3] == pytest.approx(mass * (hx * hx + hy * hy) / 3.0, rel=1e-5)


@pytest.mark.parametrize("r,c", [(0.5, (0.0, 0.0)), (0.3, (0.2, -0.1))])
def test_circle_mass_closed_form(r, c):
    _, m = po.shape_geometry(circle(r, c, density=2.0))
    mass = 2.0 * math.pi * r * r
    assert m[0] == pytest.approx(mass, rel=1e-6)
    assert m[3] == pytest.approx(mass * (0.5 * r * r + c[0] ** 2 + c[1] ** 2), rel=1e-5)


@pytest.mark.parametrize("n", [3, 5, 8])
def test_regular_polygon_hull_area_and_centroid(n):
    R = 0.8
    ang = 2 * np.pi * np.arange(n) / n + 0.1
    pts = np.stack([R * np.cos(ang) + 3.0, R * np.sin(ang) - 2.0], 1)
    rng = np.random.default_rng(n)
    g, m = po.shape_geometry(polygon(pts[rng.permutation(n)], density=1.0))
    assert int(g[0]) == n
    area = 0.5 * n * R * R * math.sin(2 * math.pi / n)
    assert m[0] == pytest.approx(area, rel=1e-5)
    assert g[2] == pytest.approx(3.0, abs=1e-5) and g[3] == pytest.approx(-2.0, abs=1e-5)
    verts = g[4:4 + 2 * n].reshape(n, 2)
    normals = g[20:20 + 2 * n].reshape(n, 2)
    # counter-clockwise hull with outward unit normals
    for i in range(n):
        e = verts[(i + 1) % n] - verts[i]
        assert e[0] * normals[i][1] - e[1] * normals[i][0] < 0
        assert np.linalg.norm(normals[i]) == pytest.approx(1.0, abs=1e-6)


def test_box_on_box_manifold_closed_form():
    m = po.collide(box(1, 1), (0, 0, 0), box(0.5, 0.5), (0.3, 1.49, 0))
    assert m["point_count"] == 2 and m["type"] == 1  # e_faceA, reference face = top of A
    pts = m["points"].reshape(2, 2)
    assert sorted(np.round(pts[:, 0], 5)) == [-0.2, 0.8]
    np.testing.assert_allclose(m["normal"], [0, 1], atol=1e-6)
    np.testing.assert_allclose(m["separations"], [-0.03, -0.03], atol=1e-6)
    # feature ids: face 2 of A against vertices 0 and 1 of B
    assert sorted(int(k) for k in m["id"]) == sorted([2 | (0 << 8) | (1 << 16), 2 | (1 << 8) | (1 << 16)])


def test_circle_circle_closed_form():
    m = po.collide(circle(1.0), (0, 0, 0), circle(0.7), (1.5, 0, 0))
    assert m["point_count"] == 1 and m["type"] == 0
    np.testing.assert_allclose(m["normal"], [1, 0], atol=1e-7)
    assert m["separations"][0] == pytest.approx(-0.2, abs=1e-6)
    np.testing.assert_allclose(m["points"][:2], [0.9, 0.0], atol=1e-6)
    assert po.collide(circle(1.0), (0, 0, 0), circle(0.7), (1.8, 0, 0))["point_count"] == 0


def test_polygon_circle_face_and_vertex_regions():
    m = po.collide(box(1, 1), (0, 0, 0), circle(0.5), (0.0, 1.3, 0))
    assert m["point_count"] == 1
    np.testing.assert_allclose(m["normal"], [0, 1], atol=1e-6)
    assert m["separations"][0] == pytest.approx(-0.21, abs=1e-6)
    m = po.collide(box(1, 1), (0, 0, 0), circle(0.5), (1.3, 1.3, 0))
    assert m["point_count"] == 1
    np.testing.assert_allclose(m["normal"], [math.sqrt(0.5)] * 2, atol=1e-6)
    assert m["separations"][0] == pytest.approx(math.hypot(0.3, 0.3) - 0.51, abs=1e-6)
    # circle first: b2Contact.Create swaps to polygon-circle
    assert po.collide(circle(0.5), (1.3, 1.3, 0), box(1, 1), (0, 0, 0))["swapped"] == 1


def test_edge_polygon_and_edge_circle():
    m = po.collide(edge(-10, 0, 10, 0), (0, 0, 0), box(0.5, 0.5), (0.0, 0.49, 0))
    assert m["point_count"] == 2
    np.testing.assert_allclose(m["normal"], [0, 1], atol=1e-6)
    np.testing.assert_allclose(m["separations"], [-0.03, -0.03], atol=1e-6)
    m = po.collide(edge(-1, 0, 1, 0), (0, 0, 0), circle(0.5), (0.0, 0.4, 0))
    assert m["point_count"] == 1 and m["type"] == 1
    assert m["separations"][0] == pytest.approx(-0.11, abs=1e-6)
    m = po.collide(edge(-1, 0, 1, 0), (0, 0, 0), circle(0.5), (1.2, 0.2, 0))
    assert m["point_count"] == 1 and m["type"] == 0 and int(m["id"][0]) == 1  # vertex B region
    # polygon given first is swapped to edge-polygon; edge-edge has no contact type
    assert po.collide(box(0.5, 0.5), (0, 0.49, 0), edge(-10, 0, 10, 0), (0, 0, 0))["swapped"] == 1
    assert po.collide(edge(-1, 0, 1, 0), (0, 0, 0), edge(-1, 0, 1, 0), (0, 0, 0)) is None


def test_polycollision_scene_is_separated():
    # examples/demo/tests/polycollision.d:38-49
    m = po.collide(box(0.2, 0.4), (0, 0, 0), box(0.5, 0.5), (19.345284, 1.5632932, 1.9160721))
    assert m["point_count"] == 0


def test_polygon_swap_symmetry():
    # Exchanging A and B mirrors the manifold: same points, same separations, negated
    # normal, and feature ids with their A and B halves exchanged.
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(400):
        a = box(float(rng.uniform(0.3, 1)), float(rng.uniform(0.3, 1)))
        b = polygon(np.stack([np.cos(t) for t in [rng.uniform(0, 6.28, 5)]] + [np.sin(rng.uniform(0, 6.28, 5))], 1) * 0.7)
        xa = (0.0, 0.0, float(rng.uniform(-3, 3)))
        xb = (float(rng.uniform(-1.5, 1.5)), float(rng.uniform(-1.5, 1.5)), float(rng.uniform(-3, 3)))
        m1 = po.collide(a, xa, b, xb)
        m2 = po.collide(b, xb, a, xa)
        if m1 is None or m1["point_count"] == 0:
            assert m2["point_count"] == 0 or m2["separations"][: m2["point_count"]].min() > -1e-4
            continue
        if m1["type"] == m2["type"]:
            continue  # reference faces differ by the k_tol hysteresis only when separations tie
        assert m1["point_count"] == m2["point_count"]
        n = m1["point_count"]
        np.testing.assert_allclose(m1["normal"], -m2["normal"], atol=1e-5)
        np.testing.assert_allclose(np.sort(m1["separations"][:n]), np.sort(m2["separations"][:n]), atol=1e-5)
        swap = lambda k: ((k >> 8) & 0xFF) | ((k & 0xFF) << 8) | (((k >> 24) & 0xFF) << 16) | (((k >> 16) & 0xFF) << 24)
        assert sorted(int(k) for k in m1["id"][:n]) == sorted(swap(int(k)) for k in m2["id"][:n])
        checked += 1
    assert checked > 50


def _random_fat(rng, n, extent=20.0, size=2.0, snap=False):
    lo = rng.uniform(-extent, extent, (n, 2)).astype(np.float32)
    sz = rng.uniform(0.05, size, (n, 2)).astype(np.float32)
    if snap:  # exact shared edges exercise the touching-counts-as-overlap rule
        lo = np.round(lo).astype(np.float32)
        sz = np.round(sz * 2).astype(np.float32) / 2 + 0.5
    return np.concatenate([lo, lo + sz], 1).astype(np.float32)


@pytest.mark.parametrize("seed,n,snap", [(888, 128, False), (1, 500, False), (2, 300, True), (3, 1, False), (4, 0, False)])
def test_tree_pairs_equal_brute_force(seed, n, snap):
    # b2DynamicTree query + b2BroadPhase sort/unique == all fat overlaps with a moved member
    # (the invariant that examples/demo/tests/dynamictreetest.d:320-335 asserts).
    rng = np.random.default_rng(seed)
    fat = _random_fat(rng, n, snap=snap)
    for frac in (1.0, 0.3, 0.0):
        moved = (rng.random(n) < frac).astype(np.uint8)
        got = po.tree_pairs(fat, moved)
        want = po.brute_force_pairs(fat, moved)
        assert got.shape == want.shape and np.array_equal(got, want)


def test_resting_box_normal_impulse_balances_gravity():
    s = scenes.single_box_on_ground(y=0.5, hx=0.5, density=1.0)
    w = po.OracleWorld(s, allow_sleep=False)
    for _ in range(180):
        w.step()
    c = w.contacts()
    touching = c[(c["flags"] & 2) != 0]
    assert len(touching) == 1
    total = touching[0]["normal_impulse"][: touching[0]["point_count"]].sum()
    assert total == pytest.approx(1.0 * 10.0 * DT, rel=1e-3)  # m * g * dt with m = 1


def test_box_falls_asleep_after_time_to_sleep():
    s = scenes.single_box_on_ground(y=0.5, hx=0.5)
    w = po.OracleWorld(s)
    asleep_at = None
    for i in range(300):
        w.step()
        if not (w.bodies()[1]["flags"] & 2):
            asleep_at = i + 1
            break
    assert asleep_at is not None and asleep_at * DT >= 0.5


def test_zero_gravity_collision_conserves_momentum():
    s = scene_of([(2, -1.0, 0.0, 0.0, 3.0, 0.0, 0.0), (2, 1.0, 0.1, 0.3, -2.0, 0.0, 0.0)], [box(0.4, 0.4), circle(0.45)])
    s.fixtures["body"] = [0, 1]
    s.gravity = (0.0, 0.0)
    w = po.OracleWorld(s)
    b0 = w.bodies()
    m = 1.0 / b0["inv_mass"]
    p0 = np.array([(m * b0["vx"]).sum(), (m * b0["vy"]).sum()])
    for _ in range(60):
        w.step()
    b = w.bodies()
    p1 = np.array([(m * b["vx"]).sum(), (m * b["vy"]).sum()])
    assert len(w.contacts()) >= 1
    np.testing.assert_allclose(p1, p0, atol=1e-4)


def test_max_translation_clamp():
    s = scene_of([(2, 0.0, 0.0, 0.0, 1000.0, 0.0, 0.0)], [circle(0.5)])
    s.fixtures["body"] = [0]
    s.gravity = (0.0, 0.0)
    w = po.OracleWorld(s)
    w.step()
    assert w.bodies()[0]["cx"] == pytest.approx(2.0, abs=1e-5)  # b2_maxTranslation per step


def test_pyramid_settles_without_deep_penetration():
    w = po.OracleWorld(scenes.pyramid())
    assert w.dynamic_count == 210
    for _ in range(300):
        w.step()
    b = w.bodies()[1:]
    assert w.max_penetration() < 4 * 0.005
    start = scenes.pyramid().bodies[1:]
    assert np.abs(b["px"] - start["px"]).max() < 0.1
