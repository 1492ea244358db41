"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The product path (dbox_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

BODY_STATE_DTYPE = np.dtype([
    ("px", "<f4"), ("py", "<f4"), ("qs", "<f4"), ("qc", "<f4"),
    ("cx", "<f4"), ("cy", "<f4"), ("a", "<f4"), ("sleep_time", "<f4"),
    ("vx", "<f4"), ("vy", "<f4"), ("w", "<f4"), ("flags", "<i4"),
    ("c0x", "<f4"), ("c0y", "<f4"), ("a0", "<f4"), ("inv_mass", "<f4"),
])
CONTACT_DTYPE = np.dtype([
    ("fixture_a", "<i4"), ("fixture_b", "<i4"), ("flags", "<i4"), ("type", "<i4"), ("point_count", "<i4"),
    ("local_normal", "<f4", (2,)), ("local_point", "<f4", (2,)), ("point_local", "<f4", (4,)),
    ("normal_impulse", "<f4", (2,)), ("tangent_impulse", "<f4", (2,)), ("id", "<u4", (2,)),
    ("friction", "<f4"), ("restitution", "<f4"),
])
EVENT_DTYPE = np.dtype([
    ("type", "<i4"), ("fixture_a", "<i4"), ("fixture_b", "<i4"), ("point_count", "<i4"),
    ("normal_impulse", "<f4", (2,)), ("tangent_impulse", "<f4", (2,)),
])
MANIFOLD_DTYPE = np.dtype([
    ("type", "<i4"), ("point_count", "<i4"), ("swapped", "<i4"), ("pad", "<i4"),
    ("local_normal", "<f4", (2,)), ("local_point", "<f4", (2,)), ("point_local", "<f4", (4,)),
    ("id", "<u4", (2,)), ("normal", "<f4", (2,)), ("points", "<f4", (4,)), ("separations", "<f4", (2,)),
])
assert BODY_STATE_DTYPE.itemsize == 64 and CONTACT_DTYPE.itemsize == 84

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i32, f32 = C.c_void_p, C.c_int, C.c_float
        L.orc_world_create.restype = vp
        L.orc_world_create.argtypes = [f32, f32]
        L.orc_world_destroy.argtypes = [vp]
        L.orc_world_set_flags.argtypes = [vp, i32, i32, i32]
        L.orc_build.argtypes = [vp, vp, i32, vp, i32]
        L.orc_world_set_solve_order.argtypes = [vp, i32]
        L.orc_solve_log.argtypes = [vp, vp, i32]
        L.orc_step.argtypes = [vp, f32, i32, i32]
        for name in ("orc_body_count", "orc_fixture_count", "orc_contact_count", "orc_dynamic_body_count",
                     "orc_proxy_count"):
            getattr(L, name).argtypes = [vp]
        L.orc_get_bodies.argtypes = [vp, vp]
        L.orc_get_contacts.argtypes = [vp, vp, i32]
        L.orc_body_contact_pairs.argtypes = [vp, i32, vp, i32]
        L.orc_events.argtypes = [vp, vp, i32]
        L.orc_contact_world_manifold.argtypes = [vp, i32, vp]
        L.orc_get_fat_aabbs.argtypes = [vp, vp]
        L.orc_last_pairs.argtypes = [vp, vp, i32]
        L.orc_profile.argtypes = [vp, vp]
        L.orc_max_penetration.argtypes = [vp]
        L.orc_max_penetration.restype = f32
        L.orc_collide.argtypes = [vp, vp, vp, vp, vp]
        L.orc_collide_child.argtypes = [vp, i32, vp, vp, vp, vp]
        L.orc_shape_geometry.argtypes = [vp, vp, vp]
        L.orc_tree_pairs.argtypes = [vp, vp, i32, vp, i32]
        L.orc_shape_distance.argtypes = [vp, vp, vp, vp, i32, vp]
        L.orc_time_of_impact.argtypes = [vp, vp, vp, vp, f32, vp]
        L.orc_toi_events.argtypes = [vp]
        L.orc_world_set_presolve.argtypes = [vp, i32]
        L.orc_create_joint.argtypes = [vp, vp]
        L.orc_joint_count.argtypes = [vp]
        L.orc_get_joints.argtypes = [vp, vp]
        L.orc_get_joints.restype = None
        L.orc_set_rot_mode.argtypes = [i32]
        L.orc_world_set_rot_mode.argtypes = [vp, i32]
        L.orc_solve_one_constraint.argtypes = [vp, vp]
        L.orc_solve_one_constraint.restype = None
        L.orc_toi_position_once.argtypes = [vp, i32, i32, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class OracleWorld:
    """b2World restated on the CPU, built from a scene's definitions."""

    def __init__(self, scene, warm_starting=True, allow_sleep=True, colour_order=False, continuous=False,
                 exact_rot=False):
        """exact_rot: step with b2Rot.Set evaluated as the device library evaluates it
        (double-precision Cody-Waite + Taylor, rounded once) instead of glibc sinf/cosf,
        so that GPU results can be compared bit for bit.  Bodies are still created with
        glibc, as the library's host-side builder creates them."""
        L = lib()
        self.L = L
        self.h = L.orc_world_create(float(scene.gravity[0]), float(scene.gravity[1]))
        if L.orc_world_set_flags(self.h, int(warm_starting), int(allow_sleep), int(continuous)) != 0:
            raise RuntimeError("oracle flag rejected")
        L.orc_world_set_solve_order(self.h, 1 if colour_order else 0)
        self.bodies_def = np.ascontiguousarray(scene.bodies)
        self.fixtures_def = np.ascontiguousarray(scene.fixtures)
        if L.orc_build(self.h, _p(self.bodies_def), len(self.bodies_def), _p(self.fixtures_def), len(self.fixtures_def)) != 0:
            raise ValueError("oracle rejected the scene")
        L.orc_world_set_rot_mode(self.h, 1 if exact_rot else 0)
        joints = getattr(scene, "joints", None)
        if joints is not None:
            self.joints_def = np.ascontiguousarray(joints)
            for k in range(len(self.joints_def)):
                if L.orc_create_joint(self.h, _p(self.joints_def[k:k + 1])) != k:
                    raise ValueError(f"oracle rejected joint {k}")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_world_destroy(self.h)
            self.h = None

    def step(self, dt=1.0 / 60.0, vel=8, pos=3, n=1):
        for _ in range(n):
            self.L.orc_step(self.h, dt, vel, pos)

    def bodies(self) -> np.ndarray:
        out = np.zeros(self.L.orc_body_count(self.h), BODY_STATE_DTYPE)
        self.L.orc_get_bodies(self.h, _p(out))
        return out

    def contacts(self) -> np.ndarray:
        n = self.L.orc_contact_count(self.h)
        out = np.zeros(n, CONTACT_DTYPE)
        got = self.L.orc_get_contacts(self.h, _p(out), n)
        return out[:got]

    def body_contact_pairs(self, body: int) -> np.ndarray:
        """The proxy pairs of body's contact edge list (b2Body.m_contactList), head first."""
        n = self.L.orc_body_contact_pairs(self.h, body, None, 0)
        out = np.zeros((max(n, 1), 2), np.int32)
        self.L.orc_body_contact_pairs(self.h, body, _p(out), n)
        return out[:n]

    def events(self) -> np.ndarray:
        """b2ContactListener callbacks of the last step (BeginContact 1, EndContact 2,
        PostSolve 3), in call order."""
        n = self.L.orc_events(self.h, None, 0)
        out = np.zeros(max(n, 1), EVENT_DTYPE)
        self.L.orc_events(self.h, _p(out), n)
        return out[:n]

    def world_manifold(self, list_index: int) -> np.ndarray:
        out = np.zeros(1, MANIFOLD_DTYPE)
        self.L.orc_contact_world_manifold(self.h, list_index, _p(out))
        return out[0]

    def fat_aabbs(self) -> np.ndarray:
        out = np.zeros((self.L.orc_proxy_count(self.h), 4), np.float32)  # one per proxy
        self.L.orc_get_fat_aabbs(self.h, _p(out))
        return out

    def last_pairs(self) -> np.ndarray:
        n = self.L.orc_last_pairs(self.h, None, 0)
        out = np.zeros((max(n, 1), 2), np.int32)
        self.L.orc_last_pairs(self.h, _p(out), n)
        return out[:n]

    def solve_log(self) -> np.ndarray:
        n = self.L.orc_solve_log(self.h, None, 0)
        out = np.zeros((max(n, 1), 4), np.int32)
        self.L.orc_solve_log(self.h, _p(out), n)
        return out[:n]

    def profile(self) -> np.ndarray:
        out = np.zeros(8, np.float32)
        self.L.orc_profile(self.h, _p(out))
        return out

    def max_penetration(self) -> float:
        return float(self.L.orc_max_penetration(self.h))

    def set_presolve(self, mode: int):
        """PreSolve test listener: 1 record, 2 record and apply the fixed policy."""
        if self.L.orc_world_set_presolve(self.h, mode) != 0:
            raise ValueError(mode)

    def toi_events(self) -> int:
        """TOI events (sub-step island solves) of the last step."""
        return int(self.L.orc_toi_events(self.h))

    def joints(self) -> np.ndarray:
        """Per joint: kept impulses (3), motor impulse, spring impulse, limit state, 0, 0."""
        out = np.zeros((self.L.orc_joint_count(self.h), 8), np.float32)
        if len(out):
            self.L.orc_get_joints(self.h, _p(out))
        return out

    def toi_position_once(self, list_index: int, toi_body: int):
        """One TOI position pass over one contact with toi_body as the TOI body:
        (c.x, c.y, a, minimum separation)."""
        out = np.zeros(4, np.float32)
        if self.L.orc_toi_position_once(self.h, int(list_index), int(toi_body), _p(out)) != 0:
            raise ValueError("no such contact or body")
        return out

    @property
    def dynamic_count(self) -> int:
        return self.L.orc_dynamic_body_count(self.h)


class rot_mode:
    """Context manager for the process-wide b2Rot.Set mode of the stand-alone calls
    (collide, distance, time of impact): 1 = the device library's evaluation."""

    def __init__(self, mode: int):
        self.mode = mode

    def __enter__(self):
        self.prev = lib().orc_set_rot_mode(int(self.mode))
        return self

    def __exit__(self, *exc):
        lib().orc_set_rot_mode(self.prev)
        return False


def solve_one_constraint(va, vb, masses, normal, ra, rb, impulses, bias, friction=0.0, tangent_speed=0.0,
                         point_count=2, iterations=1):
    """SolveVelocityConstraints on one constructed constraint between two bodies.
    va, vb: (vx, vy, w); masses: (invMassA, invIA, invMassB, invIB); ra, rb: 2x2 (point, xy);
    impulses: (n1, t1, n2, t2); bias: velocity biases of the two points.
    Returns (vA(3), vB(3), impulses(4), point count after the condition check)."""
    x = np.zeros(30, np.float32)
    x[0:3] = va
    x[3:6] = vb
    x[6:10] = masses
    x[10:12] = normal
    ra = np.asarray(ra, np.float32).reshape(2, 2)
    rb = np.asarray(rb, np.float32).reshape(2, 2)
    x[12:14], x[14:16], x[16:18], x[18:20] = ra[0], rb[0], ra[1], rb[1]
    x[20:24] = impulses
    x[24:26] = bias
    x[26] = friction
    x[27] = tangent_speed
    x[28] = point_count
    x[29] = iterations
    out = np.zeros(11, np.float32)
    lib().orc_solve_one_constraint(_p(x), _p(out))
    return out[0:3].copy(), out[3:6].copy(), out[6:10].copy(), int(out[10])


def collide(fix_a, xfa, fix_b, xfb) -> np.ndarray | None:
    """One b2Collide* call on two shapes built from fixture defs."""
    L = lib()
    a = np.ascontiguousarray(np.array(fix_a).reshape(1))
    b = np.ascontiguousarray(np.array(fix_b).reshape(1))
    xa = np.asarray(xfa, np.float32)
    xb = np.asarray(xfb, np.float32)
    out = np.zeros(1, MANIFOLD_DTYPE)
    if L.orc_collide(_p(a), _p(xa), _p(b), _p(xb), _p(out)) != 0:
        return None
    return out[0]


def collide_child(chain, child, xfa, fix_b, xfb):
    """One chain segment collided with a circle or polygon (the chain is shape A)."""
    L = lib()
    a = np.ascontiguousarray(np.array(chain).reshape(1))
    b = np.ascontiguousarray(np.array(fix_b).reshape(1))
    xa = np.asarray(xfa, np.float32)
    xb = np.asarray(xfb, np.float32)
    out = np.zeros(1, MANIFOLD_DTYPE)
    if L.orc_collide_child(_p(a), int(child), _p(xa), _p(b), _p(xb), _p(out)) != 0:
        return None
    return out[0]


def shape_distance(fix_a, xfa, fix_b, xfb, use_radii=False):
    """b2Distance with an empty simplex cache: (pointA, pointB, distance, iterations)."""
    L = lib()
    a = np.ascontiguousarray(np.array(fix_a).reshape(1))
    b = np.ascontiguousarray(np.array(fix_b).reshape(1))
    xa = np.asarray(xfa, np.float32)
    xb = np.asarray(xfb, np.float32)
    out = np.zeros(6, np.float32)
    L.orc_shape_distance(_p(a), _p(xa), _p(b), _p(xb), int(use_radii), _p(out))
    return out[0:2].copy(), out[2:4].copy(), float(out[4]), int(out[5])


TOI_FAILED, TOI_OVERLAPPED, TOI_TOUCHING, TOI_SEPARATED = 1, 2, 3, 4


def time_of_impact(fix_a, sweep_a, fix_b, sweep_b, t_max=1.0):
    """b2TimeOfImpact; sweeps are 9 floats (localCenter.xy, c0.xy, c.xy, a0, a, alpha0).
    Returns (state, t)."""
    L = lib()
    a = np.ascontiguousarray(np.array(fix_a).reshape(1))
    b = np.ascontiguousarray(np.array(fix_b).reshape(1))
    sa = np.asarray(sweep_a, np.float32)
    sb = np.asarray(sweep_b, np.float32)
    t = np.zeros(1, np.float32)
    st = L.orc_time_of_impact(_p(a), _p(sa), _p(b), _p(sb), float(t_max), _p(t))
    return int(st), float(t[0])


def shape_geometry(fix):
    L = lib()
    a = np.ascontiguousarray(np.array(fix).reshape(1))
    g = np.zeros(36, np.float32)
    m = np.zeros(4, np.float32)
    L.orc_shape_geometry(_p(a), _p(g), _p(m))
    return g, m


def tree_pairs(fat: np.ndarray, moved: np.ndarray) -> np.ndarray:
    L = lib()
    fat = np.ascontiguousarray(fat, np.float32)
    moved = np.ascontiguousarray(moved, np.uint8)
    n = len(fat)
    cap = max(16, 64 * n)
    out = np.zeros((cap, 2), np.int32)
    cnt = L.orc_tree_pairs(_p(fat), _p(moved), n, _p(out), cap)
    assert cnt <= cap
    return out[:cnt]


def brute_force_pairs(fat: np.ndarray, moved: np.ndarray) -> np.ndarray:
    """Sorted unique (i < j) pairs with touching-inclusive fat overlap and a moved member."""
    fat = np.asarray(fat, np.float32)
    lo_x, lo_y, hi_x, hi_y = fat[:, 0], fat[:, 1], fat[:, 2], fat[:, 3]
    ov = ~((lo_x[None, :] - hi_x[:, None] > 0) | (lo_y[None, :] - hi_y[:, None] > 0) |
           (lo_x[:, None] - hi_x[None, :] > 0) | (lo_y[:, None] - hi_y[None, :] > 0))
    m = np.asarray(moved, bool)
    ov &= (m[:, None] | m[None, :])
    i, j = np.nonzero(np.triu(ov, 1))
    return np.stack([i, j], 1).astype(np.int32)
