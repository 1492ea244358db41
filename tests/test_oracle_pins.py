"""Pins of oracle branches that the scene tests leave open, each against something
the oracle does not compute itself: a brute-force LCP, closed-form motion, a
brute-force filtered pair set, and the device library's sin/cos method.

Every check is chosen so that a plausible mistake in the restatement fails it:
  block solver (b2contactsolver.d:633-764)      a wrong case formula, a dropped
      K*a term or a wrong point's mass picks the wrong active set or impulse;
  friction (b2contactsolver.d:505-531)          a clamp to the new instead of the old
      normal impulse, or a tangent-speed sign error, moves the sliding speed;
  restitution (b2contactsolver.d:408-414)       a bias sign error, a threshold on |vRel|
      instead of vRel < -threshold, or min instead of max mixing changes the bounce;
  Pade damping (b2island.d:108-109)             exp(-c h) instead of 1/(1 + c h) drifts
      by 2e-3 relative in 60 steps;
  contact filter (b2worldcallbacks.d:52-65)     dropping the group rule or testing only
      one side's mask changes the contact set;
  b2_toiBaugarte (b2contactsolver.d:213)        0.2 instead of 0.75 moves a circle by
      5 mm instead of 18.75 mm in one TOI position pass;
  dtRatio (b2world.d:394, b2contactsolver.d:309-318)  a warm start that is not scaled by
      dt / dt0 leaves a resting box with a velocity after a dt change.
"""
import math

import numpy as np
import pytest

from dbox_amd import scenes
from oracle import pyoracle as po


# ---------------------------------------------------------------------------
# block solver vs brute-force LCP
# ---------------------------------------------------------------------------
def _cross(a, b):
    return a[0] * b[1] - a[1] * b[0]


def _lcp_cases(K, b):
    """All index sets of the 2x2 LCP  w = K x + b, x >= 0, w >= 0, x.w = 0 that hold,
    with their solution, by enumeration (float64)."""
    out = []
    det = K[0, 0] * K[1, 1] - K[0, 1] * K[1, 0]
    x = -np.array([K[1, 1] * b[0] - K[0, 1] * b[1], -K[1, 0] * b[0] + K[0, 0] * b[1]]) / det
    if x[0] >= 0 and x[1] >= 0:
        out.append((1, x))
    x1 = -b[0] / K[0, 0]
    if x1 >= 0 and K[1, 0] * x1 + b[1] >= 0:
        out.append((2, np.array([x1, 0.0])))
    x2 = -b[1] / K[1, 1]
    if x2 >= 0 and K[0, 1] * x2 + b[0] >= 0:
        out.append((3, np.array([0.0, x2])))
    if b[0] >= 0 and b[1] >= 0:
        out.append((4, np.zeros(2)))
    return out


def _random_constraint(rng):
    m = rng.uniform(0.2, 2.0, 2)
    inv_i = rng.uniform(0.2, 4.0, 2)
    ang = rng.uniform(0, 2 * math.pi)
    n = np.array([math.cos(ang), math.sin(ang)])
    t = np.array([n[1], -n[0]])
    c = rng.uniform(-0.5, 0.5, 2)
    half = rng.uniform(0.2, 0.8)
    ra = np.array([c + half * t, c - half * t])
    rb = ra + rng.uniform(-1.0, 1.0, 2)
    va = rng.uniform(-2, 2, 3)
    vb = rng.uniform(-2, 2, 3)
    a = rng.uniform(0, 1.0, 2) * (rng.uniform(0, 1, 2) < 0.7)
    bias = rng.uniform(0, 1.0, 2) * (rng.uniform(0, 1, 2) < 0.5)
    return m, inv_i, n.astype(np.float32), ra.astype(np.float32), rb.astype(np.float32), va, vb, a, bias


def _normal_speeds(va, vb, n, ra, rb):
    out = []
    for j in range(2):
        dvx = vb[0] - vb[2] * rb[j][1] - va[0] + va[2] * ra[j][1]
        dvy = vb[1] + vb[2] * rb[j][0] - va[1] - va[2] * ra[j][0]
        out.append(dvx * n[0] + dvy * n[1])
    return np.array(out)


def test_block_solver_matches_brute_force_lcp():
    """The two-point normal block solver returns the unique solution of its 2x2 LCP
    (K is positive definite), found by enumerating the four index sets in float64,
    and leaves velocities that satisfy complementarity; all four cases occur."""
    rng = np.random.default_rng(7)
    seen = {1: 0, 2: 0, 3: 0, 4: 0}
    checked = 0
    for _ in range(3000):
        m, inv_i, n, ra, rb, va, vb, a, bias = _random_constraint(rng)
        masses = (m[0], inv_i[0], m[1], inv_i[1])
        va32, vb32 = va.astype(np.float32), vb.astype(np.float32)
        oa, ob, imp, pc = po.solve_one_constraint(va32, vb32, masses, n, ra, rb, (a[0], 0, a[1], 0), bias)
        if pc != 2:
            continue
        n64, ra64, rb64 = n.astype(np.float64), ra.astype(np.float64), rb.astype(np.float64)
        mm = np.float32(masses).astype(np.float64)
        rn_a = [_cross(ra64[j], n64) for j in range(2)]
        rn_b = [_cross(rb64[j], n64) for j in range(2)]
        K = np.array([[mm[0] + mm[2] + mm[1] * rn_a[i] * rn_a[j] + mm[3] * rn_b[i] * rn_b[j] for j in range(2)]
                      for i in range(2)])
        a32 = np.float32(a).astype(np.float64)
        vn = _normal_speeds(va32.astype(np.float64), vb32.astype(np.float64), n64, ra64, rb64)
        b = vn - np.float32(bias).astype(np.float64) - K @ a32
        cases = _lcp_cases(K, b)
        # skip samples within float noise of a case boundary
        margins = np.abs(np.concatenate([b, -b / np.diag(K), np.linalg.solve(K, -b)]))
        if len(cases) != 1 or margins.min() < 1e-4:
            continue
        case, x = cases[0]
        seen[case] += 1
        checked += 1
        got = imp[[0, 2]].astype(np.float64)
        assert np.allclose(got, x, rtol=2e-4, atol=2e-5), (case, got, x)
        # complementarity on the velocities the solver left
        vn_after = _normal_speeds(oa.astype(np.float64), ob.astype(np.float64), n64, ra64, rb64) - bias
        assert np.all(vn_after >= -2e-4), vn_after
        assert np.all(np.abs(got * vn_after) <= 2e-4 * (1.0 + np.abs(got))), (got, vn_after)
    assert checked > 1500
    assert min(seen.values()) > 50, seen


def test_block_solver_downgrades_coincident_points():
    """Two points at the same place make K singular: the condition check keeps one
    point (b2contactsolver.d:435-447) and only point 0's impulse changes."""
    n = np.array([0.0, 1.0], np.float32)
    ra = np.array([[0.3, -0.5], [0.3, -0.5]], np.float32)
    rb = np.array([[0.3, 0.5], [0.3, 0.5]], np.float32)
    oa, ob, imp, pc = po.solve_one_constraint((0, 0, 0), (0, -1.0, 0), (0, 0, 1.0, 0.5), n, ra, rb,
                                              (0.1, 0, 0.25, 0), (0, 0))
    assert pc == 1
    assert imp[2] == np.float32(0.25)
    # one point: lambda = -normalMass * vn, normalMass = 1 / (mB + iB rnB^2)
    rnb = 0.3 * 1.0 - 0.5 * 0.0
    k = 1.0 + 0.5 * rnb * rnb
    vn = -1.0 + 0.0  # body B moves down, A is fixed
    expect = max(0.1 - vn / k, 0.0)
    assert imp[0] == pytest.approx(expect, rel=1e-6)


def test_friction_sticks_or_slides_at_the_cone():
    """One point under the bodies' centres (r parallel to the normal), so the normal
    impulse does not change the sliding speed: inside the cone the point's relative
    tangential speed ends at the contact's tangent speed; at the cone the tangent
    impulse is friction times the normal impulse from before this iteration."""
    n = np.array([0.0, 1.0], np.float32)
    ra = np.array([[0.0, 0.5], [0, 0]], np.float32)
    rb = np.array([[0.0, -0.5], [0, 0]], np.float32)
    for vx, ts, friction, old_n in [(0.3, 0.0, 0.6, 2.0), (0.3, 0.1, 0.6, 2.0), (-0.2, -0.05, 0.6, 2.0),
                                    (5.0, 0.0, 0.2, 0.5), (-5.0, 0.3, 0.2, 0.5)]:
        oa, ob, imp, _ = po.solve_one_constraint((0, 0, 0), (vx, -0.1, 0), (0, 0, 1.0, 0.0), n, ra, rb,
                                                 (old_n, 0.0, 0, 0), (0, 0), friction=friction,
                                                 tangent_speed=ts, point_count=1)
        tangent = np.array([n[1], -n[0]])  # b2Cross(normal, 1)
        vt = float(ob[0] * tangent[0] + ob[1] * tangent[1])
        cone = friction * old_n
        if abs(imp[1]) < cone * (1 - 1e-6):
            assert vt == pytest.approx(ts, abs=1e-6)
        else:
            assert abs(imp[1]) == pytest.approx(cone, rel=1e-6)
            assert abs(vt - ts) > 0.1
        assert imp[0] >= 0.0


# ---------------------------------------------------------------------------
# restitution, damping, filter, TOI Baumgarte, dtRatio on whole worlds
# ---------------------------------------------------------------------------
def _ball_scene(vy, y, e_ball, e_ground, radius=0.5):
    b = scenes.body_defs(2)
    b[0]["type"] = scenes.STATIC
    b[1]["type"] = scenes.DYNAMIC
    b[1]["py"] = y
    b[1]["vy"] = vy
    f = scenes.fixture_defs(2)
    scenes._edge(f[0], 0, -10.0, 0.0, 10.0, 0.0, 0.0)
    f[0]["restitution"] = e_ground
    f[1]["body"] = 1
    f[1]["shape"] = scenes.SHAPE_CIRCLE
    f[1]["radius"] = radius
    f[1]["density"] = 1.0
    f[1]["friction"] = 0.0
    f[1]["restitution"] = e_ball
    return scenes.Scene(b, f)


@pytest.mark.parametrize("vy,e_ball,e_ground", [(-6.0, 0.5, 0.0), (-6.0, 0.1, 0.8), (-4.0, 1.0, 0.0)])
def test_restitution_bounce_closed_form(vy, e_ball, e_ground):
    """A ball meeting the ground faster than b2_velocityThreshold leaves it at
    e = max(e_ball, e_ground) times its approach speed within the same step
    (vn after the solve = -e * vRel, vRel taken after gravity is integrated)."""
    dt = 1.0 / 60.0
    w = po.OracleWorld(_ball_scene(vy, 0.6, e_ball, e_ground))
    e = max(e_ball, e_ground)
    prev = w.bodies()[1]
    for _ in range(30):
        w.step(dt)
        cur = w.bodies()[1]
        if cur["vy"] > 0:
            v_rel = float(np.float32(prev["vy"]) + np.float32(dt) * np.float32(-10.0))
            assert v_rel < -1.0
            assert cur["vy"] == pytest.approx(-e * v_rel, rel=1e-5)
            return
        prev = cur
    raise AssertionError("the ball never bounced")


def test_no_bounce_below_velocity_threshold():
    """Below b2_velocityThreshold the bias is zero: the ball stops instead of bouncing."""
    dt = 1.0 / 60.0
    w = po.OracleWorld(_ball_scene(-0.5, 0.51, 1.0, 1.0))
    for _ in range(5):
        w.step(dt)
        assert w.bodies()[1]["vy"] <= 1e-6
    assert abs(w.bodies()[1]["vy"]) < 1e-5


def test_pade_damping_closed_form():
    """A free body with linear and angular damping in zero gravity:
    v_n = v_0 (1 / (1 + h c))^n and x_n = x_0 + h sum_k v_k."""
    b = scenes.body_defs(1)
    b[0]["type"] = scenes.DYNAMIC
    b[0]["vx"], b[0]["vy"], b[0]["w"] = 3.0, -4.0, 2.0
    b[0]["linear_damping"], b[0]["angular_damping"] = 0.5, 0.3
    f = scenes.fixture_defs(1)
    f[0]["shape"] = scenes.SHAPE_BOX
    f[0]["p"][:2] = (0.5, 0.5)
    f[0]["density"] = 1.0
    w = po.OracleWorld(scenes.Scene(b, f, gravity=(0.0, 0.0)))
    h = 1.0 / 60.0
    x, vx, vy, om = 0.0, 3.0, -4.0, 2.0
    for n in range(1, 61):
        w.step(h)
        vx /= 1.0 + h * 0.5
        vy /= 1.0 + h * 0.5
        om /= 1.0 + h * 0.3
        x += h * vx
        s = w.bodies()[0]
        assert s["vx"] == pytest.approx(vx, rel=2e-6)
        assert s["vy"] == pytest.approx(vy, rel=2e-6)
        assert s["w"] == pytest.approx(om, rel=2e-6)
    assert s["cx"] == pytest.approx(x, rel=1e-5)
    assert abs(s["vx"] - 3.0 * math.exp(-0.5)) > 1e-3  # exp(-c t) would differ


def _should_collide(fa, fb):
    # b2ContactFilter.ShouldCollide (b2worldcallbacks.d:52-65)
    if fa["group"] == fb["group"] and fa["group"] != 0:
        return fa["group"] > 0
    return (int(fa["mask"]) & int(fb["category"])) != 0 and (int(fa["category"]) & int(fb["mask"])) != 0


def filter_grid(nx=12, ny=8, seed=5, gap=0.05):
    """Boxes on a grid whose fat AABBs overlap their neighbours' while the shapes do
    not touch, with seeded categories, masks and groups, in zero gravity: nothing
    moves in the first step, so its contacts are the filtered fat-overlap pairs."""
    rng = scenes.Stream(seed)
    n = nx * ny
    b = scenes.body_defs(n)
    b["type"] = scenes.DYNAMIC
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    b["px"] = (ix.ravel() * (1.0 + gap)).astype(np.float32)
    b["py"] = (iy.ravel() * (1.0 + gap)).astype(np.float32)
    f = scenes.fixture_defs(n)
    f["body"] = np.arange(n)
    f["shape"] = scenes.SHAPE_BOX
    f["p"][:, 0] = 0.5 - 0.01
    f["p"][:, 1] = 0.5 - 0.01
    f["density"] = 1.0
    cats = np.array([1, 2, 4, 8], np.uint16)
    f["category"] = cats[rng.integers(n, 0, 3)]
    f["mask"] = np.array([0xFFFF, 0xFFFF ^ 2, 0xFFFF ^ 4, 1 | 8, 0], np.uint16)[rng.integers(n, 0, 4)]
    f["group"] = np.array([0, 0, 0, -1, -2, 1, 2], np.int16)[rng.integers(n, 0, 6)]
    return scenes.Scene(b, f, gravity=(0.0, 0.0))


def test_contact_filter_matches_brute_force():
    """Contacts after a step in which nothing moves are exactly the proxy pairs whose
    fat AABBs overlap (touching counts) and that pass ShouldCollide; the scene has
    pairs of every kind (same positive / negative group, mask exclusion on either
    side)."""
    s = filter_grid()
    w = po.OracleWorld(s)
    w.step(1.0 / 60.0)
    fat = w.fat_aabbs()
    allp = po.brute_force_pairs(fat, np.ones(len(fat), bool))
    keep = {(int(i), int(j)) for i, j in allp if _should_collide(s.fixtures[i], s.fixtures[j])}
    got = {tuple(sorted((int(c["fixture_a"]), int(c["fixture_b"])))) for c in w.contacts()}
    assert got == keep
    # every rule decided some pair
    dropped = {(int(i), int(j)) for i, j in allp} - keep
    grp = [p for p in dropped if s.fixtures[p[0]]["group"] == s.fixtures[p[1]]["group"] < 0]
    pos = [p for p in keep if s.fixtures[p[0]]["group"] == s.fixtures[p[1]]["group"] > 0 and
           not ((int(s.fixtures[p[0]]["mask"]) & int(s.fixtures[p[1]]["category"])) and
                (int(s.fixtures[p[0]]["category"]) & int(s.fixtures[p[1]]["mask"])))]
    masked = [p for p in dropped if s.fixtures[p[0]]["group"] == 0]
    assert grp and pos and masked, (len(grp), len(pos), len(masked))


def test_filter_behaviour_pass_through_and_one_sided_mask():
    """A box in the same negative group as the box under it falls through it; a box in
    the same positive group lands on it even with masks that exclude each other; a box
    whose own mask excludes the other's category falls through even though the other's
    mask accepts it."""
    def two_boxes(ga, gb, cat_a=1, mask_a=0xFFFF, cat_b=1, mask_b=0xFFFF):
        b = scenes.body_defs(3)
        b[0]["type"] = scenes.STATIC
        b[1]["type"] = b[2]["type"] = scenes.DYNAMIC
        b[1]["py"], b[2]["py"] = 0.5, 2.0
        f = scenes.fixture_defs(3)
        scenes._edge(f[0], 0, -10, 0, 10, 0, 0.6)
        f[0]["category"], f[0]["mask"] = 0x8000, 0xFFFF
        for i, (g, c, m) in zip((1, 2), ((ga, cat_a, mask_a | 0x8000), (gb, cat_b, mask_b | 0x8000))):
            f[i]["body"] = i
            f[i]["shape"] = scenes.SHAPE_BOX
            f[i]["p"][:2] = (0.5, 0.5)
            f[i]["density"] = 1.0
            f[i]["group"], f[i]["category"], f[i]["mask"] = g, c, m
        w = po.OracleWorld(scenes.Scene(b, f))
        w.step(1.0 / 60.0, n=120)
        return float(w.bodies()[2]["cy"])
    # resting heights: one box 0.5 m plus the two polygon/edge skins (0.02 m, less the
    # slop the position solver leaves), two boxes 1.5 m plus two skins
    stacked, fell = (1.5, 1.55), (0.5, 0.53)
    assert stacked[0] < two_boxes(0, 0) < stacked[1]
    assert fell[0] < two_boxes(-3, -3) < fell[1]                             # same negative group
    assert stacked[0] < two_boxes(3, 3, 1, 2, 4, 8) < stacked[1]             # positive group beats masks
    assert fell[0] < two_boxes(0, 0, 1, 0xFFFF, 2, 0xFFFF ^ 1) < fell[1]     # B's mask rejects A
    assert fell[0] < two_boxes(0, 0, 1, 0xFFFF ^ 2, 2, 0xFFFF) < fell[1]     # A's mask rejects B


@pytest.mark.parametrize("sep,expect_dy", [(-0.03, 0.75 * 0.025), (-0.012, 0.75 * 0.007), (-0.5, 0.2),
                                           (-0.004, 0.0)])
def test_toi_baumgarte_closed_form(sep, expect_dy):
    """One TOI position pass on a circle resting into a static edge (r parallel to the
    normal, so K = invMass): the circle rises by -C = -clamp(0.75 (s + slop), -0.2, 0)."""
    radius = 0.5
    y = np.float32(radius + 0.01 + sep)  # separation = centre height - edge radius - circle radius
    s = _ball_scene(0.0, float(y), 0.0, 0.0, radius)
    w = po.OracleWorld(s)
    w.step(0.0)  # dt = 0: FindNewContacts and Collide only, nothing moves
    c = w.contacts()
    assert len(c) == 1 and c[0]["point_count"] == 1
    out = w.toi_position_once(0, 1)
    assert out[3] == pytest.approx(sep, abs=2e-6)
    assert out[1] - y == pytest.approx(expect_dy, abs=3e-6)
    assert out[0] == 0.0


def test_dt_ratio_scales_the_warm_start():
    """A box at rest on the ground with converged impulses m g dt0; the next step
    changes dt and runs no velocity iteration, so the warm start alone must hold the
    box: its impulses become (dt / dt0) times the old ones and its velocity stays at
    zero.  Without the scaling the box would keep g (dt0 - dt)."""
    s = scenes.single_box_on_ground(y=0.5 + 0.005, hx=0.5, density=2.0)
    w = po.OracleWorld(s, allow_sleep=False)
    dt0 = 1.0 / 60.0
    w.step(dt0, n=120)
    c0 = w.contacts()[0]
    mass = 2.0 * 1.0 * 1.0
    assert float(c0["normal_impulse"].sum()) == pytest.approx(mass * 10.0 * dt0, rel=1e-4)
    for dt in (1.0 / 30.0, 1.0 / 120.0):
        w.step(dt, vel=0, pos=3)
        c = w.contacts()[0]
        ratio = np.float32(np.float32(1.0) / np.float32(dt0)) * np.float32(dt)
        assert float(c["normal_impulse"].sum()) == pytest.approx(float(c0["normal_impulse"].sum()) * ratio, rel=1e-6)
        assert abs(float(w.bodies()[1]["vy"])) < 1e-4
        w.step(dt0, n=60)  # back to the converged state for dt0
        c0 = w.contacts()[0]


# ---------------------------------------------------------------------------
# the exact sin/cos mode
# ---------------------------------------------------------------------------
def test_exact_rot_mode_is_correctly_rounded_nearly_always():
    """Mode 1 of b2Rot.Set rounds sin and cos once from double precision: on 20,000
    seeded angles it equals the correctly rounded float of math.sin / math.cos except
    at a handful of arguments whose value lies within a few double ulps of a float
    rounding boundary, and then by one float ulp."""
    rng = np.random.default_rng(3)
    ang = np.concatenate([rng.uniform(-7, 7, 10000), rng.uniform(-2000, 2000, 10000)]).astype(np.float32)
    got = np.zeros((len(ang), 2), np.float32)
    # free bodies in zero gravity: the step's SynchronizeTransform sets m_xf.q from the
    # unchanged sweep angle
    b = scenes.body_defs(1)
    b[0]["type"] = scenes.DYNAMIC
    f = scenes.fixture_defs(1)
    f[0]["shape"] = scenes.SHAPE_CIRCLE
    f[0]["radius"] = 1.0
    f[0]["density"] = 1.0
    for k in range(0, len(ang), 2000):
        bb = np.repeat(b, 2000)
        bb["angle"] = ang[k:k + 2000]
        bb["px"] = np.arange(2000) * 10.0
        ff = np.repeat(f, 2000)
        ff["body"] = np.arange(2000)
        w = po.OracleWorld(scenes.Scene(bb, ff, gravity=(0.0, 0.0)), exact_rot=True)
        w.step(1.0 / 60.0)
        st = w.bodies()
        got[k:k + 2000, 0] = st["qs"]
        got[k:k + 2000, 1] = st["qc"]
    ref = np.stack([np.float32(np.sin(ang.astype(np.float64))), np.float32(np.cos(ang.astype(np.float64)))], 1)
    diff = got != ref
    ulps = np.abs(got.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert diff.sum() <= 10, diff.sum()
    assert ulps.max() <= 1


def test_body_edge_lists_are_the_world_list_filtered():
    # AddPair prepends a new contact to the world list and to both bodies' edge lists
    # (b2contactmanager.d:138-166); Destroy unlinks it from all three (:88-127), also
    # when SolveTOI's FindNewContacts adds pairs.  So a body's edge list is the world
    # list filtered to that body, in the same order: the rule a contact-list mirror
    # rebuilt from the world list alone relies on.  A prepend turned into an append
    # on either list, or an unlink that skips one, breaks it.
    s = scenes.projectile_room(30, seed=5)
    o = po.OracleWorld(s, continuous=True)
    fbody = s.fixtures["body"]
    checked, events = 0, 0
    for step in range(1, 61):
        o.step()
        events += o.toi_events()
        if step % 6:
            continue
        wl = o.contacts()
        pairs = np.stack([wl["fixture_a"], wl["fixture_b"]], 1) if len(wl) else np.zeros((0, 2), np.int32)
        for b in range(len(s.bodies)):
            mine = pairs[(fbody[pairs[:, 0]] == b) | (fbody[pairs[:, 1]] == b)]
            got = o.body_contact_pairs(b)
            assert np.array_equal(got, mine), (step, b)
            checked += len(got)
    assert checked > 200 and events > 0
