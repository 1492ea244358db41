"""ctypes binding of libdbox_hip.so (include/dbox_hip.h).

``StepWorlds`` holds one handle: add worlds built from scene definitions, then
call ``step`` the way dbox code calls ``b2World.Step(dt, velocityIterations,
positionIterations)`` (src/dbox/dynamics/b2world.d:367-434).  There is no CPU
fallback: if the native library cannot be loaded, the import fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import scenes

from .scenes import BODY_DTYPE, FIXTURE_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdbox_hip.so")

DBX_OK, DBX_E_INVALID_ARG, DBX_E_CAPACITY, DBX_E_UNSUPPORTED, DBX_E_HIP, DBX_E_LOCKED, DBX_E_STATE = 0, -1, -2, -3, -4, -5, -6
WARM, SLEEP, CCD, ORDERED, LEARNED, RESIDUAL, LOG, PRESOLVE = 1, 2, 4, 8, 16, 32, 64, 128
DEFAULT_FLAGS = WARM | SLEEP

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

EVENT_BEGIN, EVENT_END, EVENT_POST_SOLVE, EVENT_PRE_SOLVE = 1, 2, 3, 4
SET_ENABLED, SET_FRICTION, SET_RESTITUTION, SET_TANGENT_SPEED = 1, 2, 4, 8
SETTING_DTYPE = np.dtype([  # dbx_contact_setting
    ("fixture_a", "<i4"), ("fixture_b", "<i4"), ("mask", "<i4"), ("enabled", "<i4"),
    ("friction", "<f4"), ("restitution", "<f4"), ("tangent_speed", "<f4"),
])
EVENT_DTYPE = np.dtype([
    ("type", "<i4"), ("fixture_a", "<i4"), ("fixture_b", "<i4"), ("point_count", "<i4"),
    ("normal_impulse", "<f4", (2,)), ("tangent_impulse", "<f4", (2,)),
])


class DbxConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("contacts_per_fixture", C.c_int32), ("block_size", C.c_int32),
                ("use_graph", C.c_int32)]


class DbxProfile(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("step", "collide", "solve", "solveInit", "solveVelocity", "solvePosition",
                                         "broadphase", "solveTOI")]


class DbxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"dbx status {status}: {msg}")
        self.status = status


_lib = None

# Every entry point declared in include/dbox_hip.h.
EXPORTS = [
    "dbx_create", "dbx_destroy", "dbx_last_error", "dbx_add_world", "dbx_add_joints", "dbx_read_joints", "dbx_finalize", "dbx_step", "dbx_sync",
    "dbx_world_count", "dbx_world_info", "dbx_read_bodies", "dbx_read_contacts", "dbx_read_events", "dbx_read_fat_aabbs",
    "dbx_read_pairs", "dbx_profile_read", "dbx_profile_totals", "dbx_phase_totals", "dbx_totals", "dbx_read_solver_stats", "dbx_broadphase_pairs",
    "dbx_collide_pairs", "dbx_broadphase_pairs_noalloc", "dbx_read_schedule", "dbx_read_tiers", "dbx_set_residual", "dbx_write_bodies",
    "dbx_autotune", "dbx_set_tier_blocks", "dbx_set_graph", "dbx_read_residuals", "dbx_read_toi_events", "dbx_write_contact_settings", "dbx_step_finish", "dbx_read_proxies",
    "dbx_mlp_init", "dbx_mlp_set", "dbx_mlp_get", "dbx_mlp_forward", "dbx_mlp_gradients", "dbx_mlp_log_enable",
    "dbx_mlp_log_count", "dbx_mlp_log_read", "dbx_mlp_log_clear", "dbx_mlp_fit_normalization", "dbx_mlp_train",
]

MLP_PARAMS = 5250
MLP_NORM = 32
MLP_FEATURES = 14


def lib():
    """Load the native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        vp, i32, f32, u32 = C.c_void_p, C.c_int32, C.c_float, C.c_uint32
        L.dbx_create.argtypes = [vp, C.POINTER(vp)]
        L.dbx_destroy.argtypes = [vp]
        L.dbx_destroy.restype = None
        L.dbx_last_error.argtypes = [vp]
        L.dbx_last_error.restype = C.c_char_p
        L.dbx_add_world.argtypes = [vp, vp, i32, vp, i32, f32, f32, C.POINTER(i32)]
        L.dbx_finalize.argtypes = [vp]
        L.dbx_add_joints.argtypes = [vp, i32, vp, i32]
        L.dbx_read_joints.argtypes = [vp, i32, vp, i32]
        L.dbx_step.argtypes = [vp, f32, i32, i32, u32]
        L.dbx_sync.argtypes = [vp]
        L.dbx_world_count.argtypes = [vp, C.POINTER(i32)]
        L.dbx_world_info.argtypes = [vp, i32, vp]
        L.dbx_read_bodies.argtypes = [vp, i32, vp, i32]
        L.dbx_read_contacts.argtypes = [vp, i32, vp, i32, C.POINTER(i32)]
        L.dbx_read_events.argtypes = [vp, i32, vp, i32, C.POINTER(i32)]
        L.dbx_read_fat_aabbs.argtypes = [vp, i32, vp, i32]
        L.dbx_read_pairs.argtypes = [vp, i32, vp, i32, C.POINTER(i32)]
        L.dbx_profile_read.argtypes = [vp, C.POINTER(DbxProfile)]
        L.dbx_profile_totals.argtypes = [vp, C.POINTER(DbxProfile), C.POINTER(i32), i32]
        L.dbx_totals.argtypes = [vp, vp]
        L.dbx_phase_totals.argtypes = [vp, vp, C.POINTER(i32), i32]
        L.dbx_read_solver_stats.argtypes = [vp, vp, vp, i32]
        L.dbx_broadphase_pairs.argtypes = [vp, vp, vp, i32, vp, i32, C.POINTER(i32)]
        L.dbx_collide_pairs.argtypes = [vp, vp, vp, vp, vp, i32, vp]
        L.dbx_broadphase_pairs_noalloc.argtypes = [vp, vp, vp, i32, vp, i32, vp]
        L.dbx_read_schedule.argtypes = [vp, i32, vp, i32, C.POINTER(i32)]
        L.dbx_read_tiers.argtypes = [vp, vp, i32, vp, C.POINTER(i32)]
        i64, u64 = C.c_int64, C.c_uint64
        L.dbx_set_residual.argtypes = [vp, f32, i32]
        L.dbx_write_bodies.argtypes = [vp, i32, vp, vp, i32]
        L.dbx_autotune.argtypes = [vp, i32]
        L.dbx_set_tier_blocks.argtypes = [vp, vp, i32]
        L.dbx_set_graph.argtypes = [vp, i32]
        L.dbx_read_residuals.argtypes = [vp, vp, i32]
        L.dbx_read_toi_events.argtypes = [vp, vp, i32]
        L.dbx_write_contact_settings.argtypes = [vp, i32, vp, i32]
        L.dbx_step_finish.argtypes = [vp]
        L.dbx_read_proxies.argtypes = [vp, i32, vp, vp, i32, C.POINTER(i32)]
        L.dbx_mlp_init.argtypes = [vp, u64]
        L.dbx_mlp_set.argtypes = [vp, vp, vp]
        L.dbx_mlp_get.argtypes = [vp, vp, vp]
        L.dbx_mlp_forward.argtypes = [vp, vp, i32, vp]
        L.dbx_mlp_gradients.argtypes = [vp, vp, vp, i32, vp, vp]
        L.dbx_mlp_log_enable.argtypes = [vp, i64]
        L.dbx_mlp_log_count.argtypes = [vp, C.POINTER(i64)]
        L.dbx_mlp_log_read.argtypes = [vp, i64, i64, vp, vp]
        L.dbx_mlp_log_clear.argtypes = [vp]
        L.dbx_mlp_fit_normalization.argtypes = [vp]
        L.dbx_mlp_train.argtypes = [vp, i32, i32, f32, u64, vp]
        for name in EXPORTS:
            fn = getattr(L, name)
            if name not in ("dbx_destroy", "dbx_last_error"):
                fn.restype = C.c_int32
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class StepWorlds:
    """One device handle holding a batch of independent worlds."""

    def __init__(self, device=0, contacts_per_fixture=0, block_size=0, use_graph=False):
        L = lib()
        self.L = L
        cfg = DbxConfig(device, contacts_per_fixture, block_size, int(use_graph))
        h = C.c_void_p()
        st = L.dbx_create(C.byref(cfg), C.byref(h))
        if st != DBX_OK:
            raise DbxError(st, "dbx_create failed (no HIP device or bad configuration)")
        self.h = h
        self.body_counts: list[int] = []
        self.fixture_counts: list[int] = []
        self.dynamic_counts: list[int] = []
        self.joint_counts: list[int] = []

    def close(self):
        if getattr(self, "h", None):
            self.L.dbx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, st):
        if st != DBX_OK:
            raise DbxError(st, self.L.dbx_last_error(self.h).decode())

    def add_scene(self, scene) -> int:
        b = np.ascontiguousarray(scene.bodies, BODY_DTYPE)
        f = np.ascontiguousarray(scene.fixtures, FIXTURE_DTYPE)
        wid = C.c_int32()
        self._check(self.L.dbx_add_world(self.h, _p(b), len(b), _p(f), len(f), float(scene.gravity[0]),
                                         float(scene.gravity[1]), C.byref(wid)))
        self.body_counts.append(len(b))
        self.fixture_counts.append(scenes.proxy_count(f))  # read-backs are per proxy
        self.dynamic_counts.append(int(np.count_nonzero(b["type"] == 2)))
        joints = getattr(scene, "joints", None)
        self.joint_counts.append(0 if joints is None else len(joints))
        if joints is not None and len(joints):
            j = np.ascontiguousarray(joints, scenes.JOINT_DTYPE)
            self._check(self.L.dbx_add_joints(self.h, wid.value, _p(j), len(j)))
        return wid.value

    def joints(self, world=0) -> np.ndarray:
        """Per joint: kept impulses (3), motor impulse, spring impulse, limit state, 0, 0."""
        n = self.joint_counts[world]
        out = np.zeros((n, 8), np.float32)
        self._check(self.L.dbx_read_joints(self.h, world, _p(out) if n else None, n))
        return out

    def finalize(self):
        self._check(self.L.dbx_finalize(self.h))

    def step(self, dt=1.0 / 60.0, vel=8, pos=3, flags=DEFAULT_FLAGS, n=1):
        for _ in range(n):
            self._check(self.L.dbx_step(self.h, dt, vel, pos, flags))

    def sync(self):
        self._check(self.L.dbx_sync(self.h))

    @property
    def world_count(self):
        return len(self.body_counts)

    def bodies(self, world=0) -> np.ndarray:
        out = np.zeros(self.body_counts[world], BODY_STATE_DTYPE)
        self._check(self.L.dbx_read_bodies(self.h, world, _p(out), len(out)))
        return out

    def contacts(self, world=0) -> np.ndarray:
        n = C.c_int32()
        self._check(self.L.dbx_read_contacts(self.h, world, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), CONTACT_DTYPE)
        self._check(self.L.dbx_read_contacts(self.h, world, _p(out), n.value, C.byref(n)))
        return out[:n.value]

    def events(self, world=0) -> np.ndarray:
        """The b2ContactListener callbacks of the last step (dbx_read_events), in dbox's order."""
        n = C.c_int32()
        self._check(self.L.dbx_read_events(self.h, world, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), EVENT_DTYPE)
        self._check(self.L.dbx_read_events(self.h, world, _p(out), n.value, C.byref(n)))
        return out[:n.value]

    def fat_aabbs(self, world=0) -> np.ndarray:
        out = np.zeros((self.fixture_counts[world], 4), np.float32)
        self._check(self.L.dbx_read_fat_aabbs(self.h, world, _p(out), len(out)))
        return out

    def pairs(self, world=0) -> np.ndarray:
        n = C.c_int32()
        self._check(self.L.dbx_read_pairs(self.h, world, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 2), np.int32)
        self._check(self.L.dbx_read_pairs(self.h, world, _p(out), n.value, C.byref(n)))
        return out[:n.value]

    def info(self, world=0):
        out = np.zeros(4, np.int32)
        self._check(self.L.dbx_world_info(self.h, world, _p(out)))
        return out

    def totals(self):
        out = np.zeros(4, np.int64)
        self._check(self.L.dbx_totals(self.h, _p(out)))
        return {"dynamic": int(out[0]), "contacts": int(out[1]), "touching": int(out[2]), "awake": int(out[3])}

    def profile(self) -> dict:
        p = DbxProfile()
        self._check(self.L.dbx_profile_read(self.h, C.byref(p)))
        return {k: getattr(p, k) for k, _ in DbxProfile._fields_}

    def schedule(self, world=0) -> np.ndarray:
        """(island, body A, body B, batch) of each constraint of the last step, in island order."""
        n = C.c_int32()
        self._check(self.L.dbx_read_schedule(self.h, world, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 4), np.int32)
        self._check(self.L.dbx_read_schedule(self.h, world, _p(out), n.value, C.byref(n)))
        return out[:n.value]

    def profile_totals(self, reset=True):
        """Per-phase milliseconds summed over the steps resolved since the last reset."""
        p = DbxProfile()
        n = C.c_int32()
        self._check(self.L.dbx_profile_totals(self.h, C.byref(p), C.byref(n), int(reset)))
        return {k: getattr(p, k) for k, _ in DbxProfile._fields_}, n.value

    def phase_totals(self, reset=True):
        """Lead-tier kernel milliseconds summed over the eager steps since the last reset:
        (collide, islands, solve, broadphase, toi), steps."""
        ms = np.zeros(5, np.float32)
        n = C.c_int32()
        self._check(self.L.dbx_phase_totals(self.h, _p(ms), C.byref(n), int(reset)))
        return dict(zip(("collide", "islands", "solve", "broadphase", "toi"), ms.astype(float))), n.value

    def write_bodies(self, world, states, forces=None):
        """Overwrite body poses, velocities, sleep times, awake bits and (optionally) force/torque."""
        st = np.ascontiguousarray(states)
        f = None if forces is None else np.ascontiguousarray(forces, np.float32).reshape(-1, 3)
        self._check(self.L.dbx_write_bodies(self.h, world, _p(st), None if f is None else _p(f), len(st)))

    # ---- residual mode and learned warm start ----
    def set_residual(self, tolerance=1e-3, max_iterations=100):
        self._check(self.L.dbx_set_residual(self.h, float(tolerance), int(max_iterations)))

    def mlp_init(self, seed=1):
        self._check(self.L.dbx_mlp_init(self.h, int(seed)))

    def mlp_set(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        n = None if norm is None else np.ascontiguousarray(norm, np.float32)
        assert p is None or p.size == MLP_PARAMS
        assert n is None or n.size == MLP_NORM
        self._check(self.L.dbx_mlp_set(self.h, None if p is None else _p(p), None if n is None else _p(n)))

    def mlp_get(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        n = np.zeros(MLP_NORM, np.float32)
        self._check(self.L.dbx_mlp_get(self.h, _p(p), _p(n)))
        return p, n

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_FEATURES)
        y = np.zeros((len(x), 2), np.float32)
        self._check(self.L.dbx_mlp_forward(self.h, _p(x), len(x), _p(y)))
        return y

    def mlp_gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.L.dbx_mlp_gradients(self.h, _p(x), _p(t), len(x), _p(g), _p(loss)))
        return g, float(loss[0])

    def mlp_log_enable(self, capacity):
        self._check(self.L.dbx_mlp_log_enable(self.h, int(capacity)))

    def mlp_log_count(self) -> int:
        c = C.c_int64()
        self._check(self.L.dbx_mlp_log_count(self.h, C.byref(c)))
        return c.value

    def mlp_log_read(self, first=0, n=None):
        if n is None:
            n = self.mlp_log_count() - first
        x = np.zeros((n, MLP_FEATURES), np.float32)
        t = np.zeros((n, 2), np.float32)
        self._check(self.L.dbx_mlp_log_read(self.h, int(first), int(n), _p(x), _p(t)))
        return x, t

    def mlp_log_clear(self):
        self._check(self.L.dbx_mlp_log_clear(self.h))

    def mlp_fit_normalization(self):
        self._check(self.L.dbx_mlp_fit_normalization(self.h))

    def mlp_train(self, steps, batch=4096, lr=1e-3, seed=1):
        loss = np.zeros(max(steps, 1), np.float32)
        self._check(self.L.dbx_mlp_train(self.h, int(steps), int(batch), float(lr), int(seed), _p(loss)))
        return loss[:steps]

    def autotune(self, steps_per_candidate=4):
        self._check(self.L.dbx_autotune(self.h, int(steps_per_candidate)))

    def set_graph(self, enabled=True):
        self._check(self.L.dbx_set_graph(self.h, int(enabled)))

    def set_tier_blocks(self, blocks):
        b = np.ascontiguousarray(blocks, np.int32)
        self._check(self.L.dbx_set_tier_blocks(self.h, _p(b), len(b)))

    def proxies(self, world=0):
        """(fixture, child) of each proxy index the read-backs use."""
        n = C.c_int32()
        self._check(self.L.dbx_read_proxies(self.h, world, None, None, 0, C.byref(n)))
        f = np.zeros(n.value, np.int32)
        c = np.zeros(n.value, np.int32)
        self._check(self.L.dbx_read_proxies(self.h, world, _p(f), _p(c), n.value, C.byref(n)))
        return f, c

    def step_finish(self):
        """The rest of a step started with the PRESOLVE flag."""
        self._check(self.L.dbx_step_finish(self.h))

    def write_contact_settings(self, world, settings):
        s = np.ascontiguousarray(settings, SETTING_DTYPE)
        self._check(self.L.dbx_write_contact_settings(self.h, world, _p(s), len(s)))

    def toi_events(self):
        """Per world: TOI events solved by the last step's continuous pass."""
        e = np.zeros(self.world_count, np.int32)
        self._check(self.L.dbx_read_toi_events(self.h, _p(e), len(e)))
        return e

    def residuals(self):
        r = np.zeros(self.world_count, np.float32)
        self._check(self.L.dbx_read_residuals(self.h, _p(r), len(r)))
        return r

    def tiers(self):
        """Launch tiers as rows (first launch slot, worlds, LDS bytes, workgroup size) and the launch order."""
        n = C.c_int32()
        rows = np.zeros((8, 4), np.int32)
        order = np.zeros(self.world_count, np.int32)
        self._check(self.L.dbx_read_tiers(self.h, _p(rows), 8, _p(order), C.byref(n)))
        return rows[:n.value], order

    def solver_stats(self):
        nw = self.world_count
        it = np.zeros(nw, np.int32)
        pen = np.zeros(nw, np.float32)
        self._check(self.L.dbx_read_solver_stats(self.h, _p(it), _p(pen), nw))
        return it, pen


def broadphase_pairs(handle: StepWorlds, fat: np.ndarray, moved: np.ndarray) -> np.ndarray:
    fat = np.ascontiguousarray(fat, np.float32)
    mv = np.ascontiguousarray(moved, np.uint8)
    n = len(fat)
    cap = max(16, 64 * n)
    out = np.zeros((cap, 2), np.int32)
    cnt = C.c_int32()
    handle._check(handle.L.dbx_broadphase_pairs(handle.h, _p(fat), _p(mv), n, _p(out), cap, C.byref(cnt)))
    return out[:cnt.value]


def collide_pairs(handle: StepWorlds, fa: np.ndarray, xa: np.ndarray, fb: np.ndarray, xb: np.ndarray) -> np.ndarray:
    fa = np.ascontiguousarray(fa, FIXTURE_DTYPE)
    fb = np.ascontiguousarray(fb, FIXTURE_DTYPE)
    xa = np.ascontiguousarray(xa, np.float32)
    xb = np.ascontiguousarray(xb, np.float32)
    n = len(fa)
    out = np.zeros((n, 32), np.float32)
    handle._check(handle.L.dbx_collide_pairs(handle.h, _p(fa), _p(xa), _p(fb), _p(xb), n, _p(out)))
    return out
