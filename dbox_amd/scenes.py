"""Seeded scene generators for the step workloads.

A scene is a pair of numpy structured arrays in creation order: body
definitions (mirroring b2BodyDef, src/dbox/dynamics/b2body.d:51-104) and
fixture definitions (b2FixtureDef, src/dbox/dynamics/b2fixture.d:49-73, plus
the shape constructor).  The byte layout is the C struct ``dbx_body_def`` /
``dbx_fixture_def`` of include/dbox_hip.h, so one array feeds both the GPU
library and any other consumer of the same definitions.

Randomness is splitmix64 seeded per workload (``0xD80C + config``) and per
world (``seed ^ world``); every draw is a 24-bit uniform in [0, 1), so the
scenes are identical on every machine.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

BODY_DTYPE = np.dtype([
    ("type", "<i4"), ("px", "<f4"), ("py", "<f4"), ("angle", "<f4"),
    ("vx", "<f4"), ("vy", "<f4"), ("w", "<f4"),
    ("linear_damping", "<f4"), ("angular_damping", "<f4"), ("gravity_scale", "<f4"),
    ("allow_sleep", "<i4"), ("awake", "<i4"), ("fixed_rotation", "<i4"), ("bullet", "<i4"),
])
FIXTURE_DTYPE = np.dtype([
    ("body", "<i4"), ("shape", "<i4"), ("count", "<i4"), ("flags", "<i4"),
    ("density", "<f4"), ("friction", "<f4"), ("restitution", "<f4"),
    ("category", "<u2"), ("mask", "<u2"), ("group", "<i2"), ("pad", "<i2"),
    ("radius", "<f4"), ("p", "<f4", (20,)),
])
# Joint definition (b2JointDef and the concrete b2*JointDef, src/dbox/dynamics/joints/):
# the byte layout of dbx_joint_def in include/dbox_hip.h.
JOINT_DTYPE = np.dtype([
    ("type", "<i4"), ("body_a", "<i4"), ("body_b", "<i4"), ("collide_connected", "<i4"),
    ("enable_limit", "<i4"), ("enable_motor", "<i4"), ("joint1", "<i4"), ("joint2", "<i4"),
    ("local_anchor_a", "<f4", (2,)), ("local_anchor_b", "<f4", (2,)), ("local_axis_a", "<f4", (2,)),
    ("reference_angle", "<f4"), ("length", "<f4"), ("frequency_hz", "<f4"), ("damping_ratio", "<f4"),
    ("lower", "<f4"), ("upper", "<f4"), ("motor_speed", "<f4"), ("max_motor", "<f4"),
    ("max_force", "<f4"), ("max_torque", "<f4"), ("linear_offset", "<f4", (2,)), ("angular_offset", "<f4"),
    ("correction_factor", "<f4"), ("ground_anchor_a", "<f4", (2,)), ("ground_anchor_b", "<f4", (2,)),
    ("length_a", "<f4"), ("length_b", "<f4"), ("ratio", "<f4"), ("target", "<f4", (2,)),
])
assert BODY_DTYPE.itemsize == 56 and FIXTURE_DTYPE.itemsize == 120 and JOINT_DTYPE.itemsize == 148
# b2JointType (b2joint.d:28-41)
REVOLUTE, PRISMATIC, DISTANCE, PULLEY, MOUSE, GEAR, WHEEL, WELD, FRICTION, ROPE, MOTOR = 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11

STATIC, KINEMATIC, DYNAMIC = 0, 1, 2
SHAPE_CIRCLE, SHAPE_EDGE, SHAPE_POLYGON, SHAPE_BOX, SHAPE_CHAIN = 0, 1, 2, 3, 4
CHAIN_PREV, CHAIN_NEXT, CHAIN_LOOP = 2, 4, 16  # chain fixture flags


def proxy_count(fixtures: np.ndarray) -> int:
    """Broadphase proxies of a fixture list: one per fixture, one per chain segment
    (b2ChainShape.GetChildCount, b2chainshape.d:129-132)."""
    f = np.asarray(fixtures)
    chain = f["shape"] == SHAPE_CHAIN
    loop = (f["flags"] & CHAIN_LOOP) != 0
    kids = np.where(chain, np.where(loop, f["count"], f["count"] - 1), 1)
    return int(kids.sum())

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class Stream:
    """splitmix64 stream; ``uniform(n)`` returns n float64 draws in [0, 1)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * _GOLDEN
            self.state = self.state + np.uint64(n) * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * math.pi * u2)

    def integers(self, n: int, lo: int, hi_inclusive: int) -> np.ndarray:
        span = hi_inclusive - lo + 1
        return lo + np.minimum((self.uniform(n) * span).astype(np.int64), span - 1)


def config_seed(config: int) -> int:
    return 0xD80C + config


@dataclass
class Scene:
    bodies: np.ndarray
    fixtures: np.ndarray
    gravity: tuple = (0.0, -10.0)
    name: str = ""
    meta: dict = field(default_factory=dict)
    joints: np.ndarray | None = None  # JOINT_DTYPE records in b2World.CreateJoint order

    @property
    def dynamic_count(self) -> int:
        return int(np.count_nonzero(self.bodies["type"] == DYNAMIC))


def body_defs(n: int) -> np.ndarray:
    b = np.zeros(n, BODY_DTYPE)
    b["gravity_scale"] = 1.0
    b["allow_sleep"] = 1
    b["awake"] = 1
    return b


def joint_defs(n: int, kind: int = 0) -> np.ndarray:
    """n joint records with the reference's definition defaults: b2DistanceJointDef
    length 1, b2PrismaticJointDef / b2WheelJointDef axis (1, 0), b2WheelJointDef
    frequency 2 Hz and damping 0.7, b2PulleyJointDef ratio 1, b2MouseJointDef max
    force 0, frequency 5 Hz and damping 0.7, b2MotorJointDef max force 1, max
    torque 1, correction factor 0.3."""
    j = np.zeros(n, JOINT_DTYPE)
    j["type"] = kind
    j["length"] = 1.0
    j["local_axis_a"] = (1.0, 0.0)
    j["ratio"] = 1.0
    if kind == WHEEL:
        j["frequency_hz"], j["damping_ratio"] = 2.0, 0.7
    if kind == MOUSE:
        j["frequency_hz"], j["damping_ratio"] = 5.0, 0.7
    if kind == MOTOR:
        j["max_force"], j["max_torque"], j["correction_factor"] = 1.0, 1.0, 0.3
    return j


def fixture_defs(n: int) -> np.ndarray:
    f = np.zeros(n, FIXTURE_DTYPE)
    f["friction"] = 0.2          # b2FixtureDef.friction default (b2fixture.d:59)
    f["category"] = 1
    f["mask"] = 0xFFFF
    return f


def _edge(f, body, x1, y1, x2, y2, friction):
    f["body"] = body
    f["shape"] = SHAPE_EDGE
    f["p"][:4] = (x1, y1, x2, y2)
    f["friction"] = friction


# --------------------------------------------------------------------------
# hello_world (examples/hello_world/hello_world.d:31-102)
# --------------------------------------------------------------------------
def hello_world() -> Scene:
    b = body_defs(2)
    b[0]["type"] = STATIC
    b[0]["py"] = -10.0
    b[1]["type"] = DYNAMIC
    b[1]["py"] = 4.0
    f = fixture_defs(2)
    f[0]["body"] = 0
    f[0]["shape"] = SHAPE_BOX
    f[0]["p"][:2] = (50.0, 10.0)
    f[1]["body"] = 1
    f[1]["shape"] = SHAPE_BOX
    f[1]["p"][:2] = (1.0, 1.0)
    f[1]["density"] = 1.0
    f[1]["friction"] = 0.3
    return Scene(b, f, name="hello_world")


# --------------------------------------------------------------------------
# Pyramid: examples/demo/tests/pyramid.d:39-78, with the workload's friction
# and a seeded +-0.01 m jitter of every box position.
# --------------------------------------------------------------------------
def pyramid(rows: int = 20, seed: int | None = None, jitter: float = 0.01, friction: float = 0.6,
            density: float = 5.0, max_boxes: int | None = None, ground_half_width: float = 40.0) -> Scene:
    seed = config_seed(1) if seed is None else seed
    rng = Stream(seed)
    xs, ys = [], []
    x = np.array([-7.0, 0.75], np.float32)
    dx = np.array([0.5625, 1.25], np.float32)
    dy = np.array([1.125, 0.0], np.float32)
    for i in range(rows):
        y = x.copy()
        for _ in range(i, rows):
            xs.append(y[0])
            ys.append(y[1])
            y = (y + dy).astype(np.float32)
        x = (x + dx).astype(np.float32)
    n = len(xs) if max_boxes is None else min(len(xs), max_boxes)
    px = np.array(xs[:n], np.float32)
    py = np.array(ys[:n], np.float32)
    if jitter:
        px = (px + rng.uniform(n, -jitter, jitter).astype(np.float32)).astype(np.float32)
        py = (py + rng.uniform(n, -jitter, jitter).astype(np.float32)).astype(np.float32)
    b = body_defs(n + 1)
    b[0]["type"] = STATIC
    b[1:]["type"] = DYNAMIC
    b[1:]["px"] = px
    b[1:]["py"] = py
    f = fixture_defs(n + 1)
    _edge(f[0], 0, -ground_half_width, 0.0, ground_half_width, 0.0, friction)
    f[1:]["body"] = np.arange(1, n + 1)
    f[1:]["shape"] = SHAPE_BOX
    f[1:]["p"][:, 0] = 0.5
    f[1:]["p"][:, 1] = 0.5
    f[1:]["density"] = density
    f[1:]["friction"] = friction
    return Scene(b, f, name=f"pyramid{rows}", meta={"rows": rows, "boxes": n})


# --------------------------------------------------------------------------
# Mixed polygons and circles in a static U container (configs 2 and 3).
# --------------------------------------------------------------------------
def mixed_container(n: int, width: float, height: float, seed: int) -> Scene:
    rng = Stream(seed)
    is_poly = rng.uniform(n) < 0.6
    nverts = rng.integers(n, 3, 8)
    prad = rng.uniform(n, 0.25, 1.0)
    crad = rng.uniform(n, 0.2, 0.8)
    jit = rng.uniform(n * 8, -0.3, 0.3).reshape(n, 8)
    density = rng.uniform(n, 0.5, 5.0)
    friction = rng.uniform(n, 0.2, 0.8)
    ux = rng.uniform(n)
    uy = rng.uniform(n)
    angle = rng.uniform(n, 0.0, 2.0 * math.pi)
    extent = np.where(is_poly, prad, crad) + 0.02
    half = 0.5 * width
    px = -half + extent + ux * (width - 2 * extent)
    py = extent + uy * (height - 2 * extent)

    b = body_defs(n + 1)
    b[0]["type"] = STATIC
    b[1:]["type"] = DYNAMIC
    b[1:]["px"] = px
    b[1:]["py"] = py
    b[1:]["angle"] = angle
    f = fixture_defs(n + 3)
    _edge(f[0], 0, -half, 0.0, half, 0.0, 0.6)
    _edge(f[1], 0, -half, 0.0, -half, height, 0.6)
    _edge(f[2], 0, half, 0.0, half, height, 0.6)
    d = f[3:]
    d["body"] = np.arange(1, n + 1)
    d["density"] = density
    d["friction"] = friction
    d["shape"] = np.where(is_poly, SHAPE_POLYGON, SHAPE_CIRCLE)
    d["radius"] = np.where(is_poly, 0.0, crad)
    d["count"] = np.where(is_poly, nverts, 0)
    k = np.arange(8)[None, :]
    ang = 2.0 * math.pi * (k + jit) / nverts[:, None]
    vx = prad[:, None] * np.cos(ang)
    vy = prad[:, None] * np.sin(ang)
    pts = np.zeros((n, 20), np.float32)
    pts[:, 0:16:2] = vx
    pts[:, 1:16:2] = vy
    valid = k < nverts[:, None]
    pts[:, 0:16:2] *= valid
    pts[:, 1:16:2] *= valid
    pts[~is_poly] = 0.0
    d["p"] = pts
    return Scene(b, f, name=f"mixed{n}", meta={"width": width, "height": height})


def big_mixed_world(n: int = 20000, seed: int | None = None) -> Scene:
    """20,000 bodies dropped into a 200 m x 400 m U container."""
    return mixed_container(n, 200.0, 400.0, config_seed(2) if seed is None else seed)


def world_sizes(worlds: int, seed: int, median: float = 64.0, sigma: float = 0.8,
                lo: int = 8, hi: int = 1024) -> np.ndarray:
    z = Stream(seed).normal(worlds)
    return np.clip(np.rint(np.exp(math.log(median) + sigma * z)), lo, hi).astype(np.int64)


def batched_worlds(worlds: int = 4096, seed: int | None = None, sizes: np.ndarray | None = None) -> list[Scene]:
    """Independent 40 m x 80 m container worlds of lognormal body counts."""
    seed = config_seed(3) if seed is None else seed
    sizes = world_sizes(worlds, seed) if sizes is None else sizes
    return [mixed_container(int(sizes[w]), 40.0, 80.0, seed ^ (w + 1) * 0x632BE5AB) for w in range(worlds)]


def big_pyramid(boxes: int = 50000, seed: int | None = None) -> Scene:
    rows = int(math.ceil((math.sqrt(8 * boxes + 1) - 1) / 2))
    return pyramid(rows=rows, seed=config_seed(4) if seed is None else seed, max_boxes=boxes,
                   ground_half_width=400.0)


def stacks_world(seed: int) -> Scene:
    """1-4 stacks of 20-60 boxes with half-extents in [0.25, 0.75] on a ground edge."""
    rng = Stream(seed)
    nstacks = int(rng.integers(1, 1, 4)[0])
    heights = rng.integers(nstacks, 20, 60)
    total = int(heights.sum())
    hx = rng.uniform(total, 0.25, 0.75).astype(np.float32)
    hy = rng.uniform(total, 0.25, 0.75).astype(np.float32)
    fr = rng.uniform(total, 0.3, 0.9)
    b = body_defs(total + 1)
    b[0]["type"] = STATIC
    f = fixture_defs(total + 1)
    _edge(f[0], 0, -60.0, 0.0, 60.0, 0.0, 0.6)
    i = 0
    for s in range(nstacks):
        x0 = -30.0 + 20.0 * s
        y = 0.0
        for _ in range(int(heights[s])):
            y += float(hy[i]) + 0.005
            bb = b[i + 1]
            bb["type"] = DYNAMIC
            bb["px"] = x0
            bb["py"] = y
            y += float(hy[i]) + 0.005
            ff = f[i + 1]
            ff["body"] = i + 1
            ff["shape"] = SHAPE_BOX
            ff["p"][:2] = (hx[i], hy[i])
            ff["density"] = 1.0
            ff["friction"] = fr[i]
            i += 1
    return Scene(b, f, name="stacks", meta={"stacks": nstacks})


def stack_worlds(worlds: int = 16384, seed: int | None = None) -> list[Scene]:
    seed = config_seed(5) if seed is None else seed
    return [stacks_world(seed ^ (w + 1) * 0x632BE5AB) for w in range(worlds)]


def single_box_on_ground(y: float = 0.5, hx: float = 0.5, density: float = 1.0, friction: float = 0.6) -> Scene:
    """One box resting on a static ground box (closed-form solver checks)."""
    b = body_defs(2)
    b[0]["type"] = STATIC
    b[1]["type"] = DYNAMIC
    b[1]["py"] = y
    f = fixture_defs(2)
    f[0]["body"] = 0
    f[0]["shape"] = SHAPE_BOX
    f[0]["p"][:5] = (20.0, 0.5, 0.0, -0.5, 0.0)
    f[0]["flags"] = 8
    f[0]["friction"] = friction
    f[1]["body"] = 1
    f[1]["shape"] = SHAPE_BOX
    f[1]["p"][:2] = (hx, hx)
    f[1]["density"] = density
    f[1]["friction"] = friction
    return Scene(b, f, name="box_on_ground")


def projectile_room(n: int = 40, seed: int | None = None, size: float = 20.0, speed: float = 150.0) -> Scene:
    """n small fast bodies (circles, boxes and polygons of 0.1-0.3 m) thrown at up
    to `speed` m/s inside a closed square room of four static edges: the workload
    of the continuous pass (b2World.SolveTOI), since without it most of them leave
    the room through its walls within a few steps."""
    rng = Stream(config_seed(6) if seed is None else seed)
    kind = rng.integers(n, 0, 2)
    r = rng.uniform(n, 0.1, 0.3)
    hx = rng.uniform(n, 0.1, 0.3)
    hy = rng.uniform(n, 0.1, 0.3)
    nverts = rng.integers(n, 3, 8)
    jit = rng.uniform(n * 8, -0.3, 0.3).reshape(n, 8)
    half = 0.5 * size
    px = rng.uniform(n, -half + 1.0, half - 1.0)
    py = rng.uniform(n, 1.0, size - 1.0)
    heading = rng.uniform(n, 0.0, 2.0 * math.pi)
    mag = rng.uniform(n, 0.5, 1.0) * speed
    w = rng.uniform(n, -10.0, 10.0)
    angle = rng.uniform(n, 0.0, 2.0 * math.pi)

    b = body_defs(n + 1)
    b[0]["type"] = STATIC
    d = b[1:]
    d["type"] = DYNAMIC
    d["px"] = px
    d["py"] = py
    d["angle"] = angle
    d["vx"] = mag * np.cos(heading)
    d["vy"] = mag * np.sin(heading)
    d["w"] = w
    f = fixture_defs(n + 4)
    _edge(f[0], 0, -half, 0.0, half, 0.0, 0.6)
    _edge(f[1], 0, -half, size, half, size, 0.6)
    _edge(f[2], 0, -half, 0.0, -half, size, 0.6)
    _edge(f[3], 0, half, 0.0, half, size, 0.6)
    g = f[4:]
    g["body"] = np.arange(1, n + 1)
    g["density"] = 1.0
    g["friction"] = 0.4
    g["shape"] = np.choose(kind, [SHAPE_CIRCLE, SHAPE_BOX, SHAPE_POLYGON])
    g["radius"] = np.where(kind == 0, r, 0.0)
    g["count"] = np.where(kind == 2, nverts, 0)
    pts = np.zeros((n, 20), np.float32)
    k = np.arange(8)[None, :]
    ang = 2.0 * math.pi * (k + jit) / nverts[:, None]
    valid = k < nverts[:, None]
    pts[:, 0:16:2] = r[:, None] * np.cos(ang) * valid
    pts[:, 1:16:2] = r[:, None] * np.sin(ang) * valid
    pts[kind == 1, 0] = hx[kind == 1]
    pts[kind == 1, 1] = hy[kind == 1]
    pts[kind == 1, 2:] = 0.0
    pts[kind == 0] = 0.0
    g["p"] = pts
    return Scene(b, f, name=f"projectiles{n}", meta={"size": size})


def chain_terrain(n: int = 60, seed: int | None = None, pieces: int = 4) -> Scene:
    """Bodies of the container workloads' shapes dropped onto a bumpy ground made of
    `pieces` chain fixtures (b2ChainShape, b2chainshape.d:24-263) of 8 vertices,
    each joined to its neighbours by ghost vertices (SetPrevVertex / SetNextVertex),
    inside a closed loop chain (CreateLoop) that bounds the world."""
    rng = Stream(config_seed(7) if seed is None else seed)
    seg = 7 * pieces
    xs = np.linspace(-20.0, 20.0, seg + 1)
    ys = rng.uniform(seg + 1, 0.0, 1.2)
    b = body_defs(n + 1)
    b[0]["type"] = STATIC
    f = fixture_defs(pieces + 1 + n)
    for k in range(pieces):
        c = f[k]
        c["body"] = 0
        c["shape"] = SHAPE_CHAIN
        c["count"] = 8
        c["friction"] = 0.6
        i0 = 7 * k
        pts = np.stack([xs[i0:i0 + 8], ys[i0:i0 + 8]], 1).astype(np.float32)
        p = np.zeros(20, np.float32)
        p[:16] = pts.ravel()
        flags = 0
        if k > 0:
            p[16:18] = (xs[i0 - 1], ys[i0 - 1])
            flags |= CHAIN_PREV
        if k < pieces - 1:
            p[18:20] = (xs[i0 + 8], ys[i0 + 8])
            flags |= CHAIN_NEXT
        c["p"] = p
        c["flags"] = flags
    box = f[pieces]
    box["body"] = 0
    box["shape"] = SHAPE_CHAIN
    box["count"] = 4
    box["flags"] = CHAIN_LOOP
    box["friction"] = 0.6
    p = np.zeros(20, np.float32)
    p[:8] = (-22.0, -2.0, 22.0, -2.0, 22.0, 40.0, -22.0, 40.0)
    box["p"] = p
    # the dynamic bodies: the mixed container's shape and material draws, above the ground
    mix = mixed_container(n, 36.0, 30.0, rng.next_u64(1)[0].item())
    b[1:] = mix.bodies[1:]
    b[1:]["py"] += 3.0
    d = f[pieces + 1:]
    d[:] = mix.fixtures[3:]
    d["body"] = np.arange(1, n + 1)
    return Scene(b, f, name=f"chains{n}")


SENSOR = 1  # fixture flags bit0: b2FixtureDef.isSensor


def sensor_container(n: int = 150, seed: int = 61, every: int = 9) -> Scene:
    """A mixed container (40 m x 80 m) with sensors: a static 16 m x 10 m sensor box
    across the middle that the falling bodies pass through, and every `every`-th
    body's fixture made a sensor (it falls through everything, reporting the
    overlaps it passes)."""
    s = mixed_container(n, 40.0, 80.0, seed)
    b = np.concatenate([s.bodies, body_defs(1)])
    b[-1]["type"] = STATIC
    b[-1]["py"] = 30.0
    f = np.concatenate([s.fixtures, fixture_defs(1)])
    z = f[-1]
    z["body"] = len(b) - 1
    z["shape"] = SHAPE_BOX
    z["p"][:2] = (8.0, 5.0)
    z["flags"] = SENSOR
    f["flags"][3::every] |= SENSOR
    return Scene(b, f, name=f"sensors{n}", meta=dict(s.meta))


# --------------------------------------------------------------------------
# Joints: a world with every supported joint kind next to contacts.
# --------------------------------------------------------------------------
def joint_world(seed: int = 1, debris: int = 12) -> Scene:
    """A 80 m wide ground edge carrying: a suspension bridge of 12 planks on revolute
    joints anchored to the ground body at both ends, with debris dropped on it; a car
    (chassis on two wheel joints with suspension springs, rear wheel motor); a chain
    of 5 balls on distance joints hanging from an anchor; a box on a rope; an L of
    three welded boxes; an elevator on a prismatic joint with limits and a motor; two
    boxes on a pulley; a box slid along the ground against a friction joint; a box
    driven to an offset by a motor joint; a box held by a mouse joint; and a revolute
    pendulum with a limit.  Positions, densities and speeds are seeded."""
    rng = Stream(config_seed(7) ^ seed)
    bodies, fixtures, joints = [], [], []

    def body(t, x, y, a=0.0, vx=0.0, vy=0.0, w=0.0):
        b = body_defs(1)[0]
        b["type"], b["px"], b["py"], b["angle"], b["vx"], b["vy"], b["w"] = t, x, y, a, vx, vy, w
        bodies.append(b)
        return len(bodies) - 1

    def box(bi, hx, hy, density=1.0, friction=0.6):
        f = fixture_defs(1)[0]
        f["body"], f["shape"], f["density"], f["friction"] = bi, SHAPE_BOX, density, friction
        f["p"][:2] = (hx, hy)
        fixtures.append(f)

    def ball(bi, r, density=1.0, friction=0.6):
        f = fixture_defs(1)[0]
        f["body"], f["shape"], f["radius"], f["density"], f["friction"] = bi, SHAPE_CIRCLE, r, density, friction
        fixtures.append(f)

    def joint(kind, a, b, **kw):
        j = joint_defs(1, kind)[0]
        j["body_a"], j["body_b"] = a, b
        for k, v in kw.items():
            j[k] = v
        joints.append(j)

    u = lambda lo, hi: float(rng.uniform(1, lo, hi)[0])
    g = body(STATIC, 0.0, 0.0)
    f = fixture_defs(1)[0]
    _edge(f, g, -40.0, 0.0, 40.0, 0.0, 0.6)
    fixtures.append(f)
    # bridge: planks 1 m long between x = -30 and -18 at y = 6
    prev, x0 = g, -30.0
    for i in range(12):
        p = body(DYNAMIC, x0 + 0.5 + i, 6.0)
        box(p, 0.5, 0.125, density=u(15.0, 25.0), friction=0.2)
        anchor = (x0 + i, 6.0)
        la = (anchor[0], anchor[1]) if prev == g else (0.5, 0.0)
        joint(REVOLUTE, prev, p, local_anchor_a=la, local_anchor_b=(-0.5, 0.0))
        prev = p
    joint(REVOLUTE, prev, g, local_anchor_a=(0.5, 0.0), local_anchor_b=(x0 + 12.0, 6.0))
    for i in range(debris):
        d = body(DYNAMIC, u(-29.0, -19.0), u(8.0, 14.0), u(0.0, 6.28))
        (box(d, u(0.2, 0.5), u(0.2, 0.5)) if i % 2 else ball(d, u(0.2, 0.5)))
    # car
    ch = body(DYNAMIC, -8.0, 1.2)
    box(ch, 1.5, 0.3, density=1.0)
    for k, xw in enumerate((-9.0, -7.0)):
        wh = body(DYNAMIC, xw, 0.45)
        ball(wh, 0.4, friction=0.9)
        joint(WHEEL, ch, wh, local_anchor_a=(xw + 8.0, -0.75), local_axis_a=(0.0, 1.0), frequency_hz=4.0,
              damping_ratio=0.7, enable_motor=1 if k == 0 else 0, motor_speed=-u(4.0, 12.0), max_motor=20.0)
    # chain of balls on distance joints from an anchor at (0, 12)
    prev, py = g, 12.0
    for i in range(5):
        bb = body(DYNAMIC, u(-0.3, 0.3) + 0.8 * (i + 1), 12.0 - 0.2 * (i + 1))
        ball(bb, 0.2)
        if prev == g:
            joint(DISTANCE, g, bb, local_anchor_a=(0.0, 12.0), length=0.8)
        else:
            joint(DISTANCE, prev, bb, length=0.8, frequency_hz=0.0 if i % 2 else 6.0, damping_ratio=0.5)
        prev = bb
    # box on a rope from (6, 12)
    rb = body(DYNAMIC, 8.0, 11.0)
    box(rb, 0.4, 0.4)
    joint(ROPE, g, rb, local_anchor_a=(6.0, 12.0), length=3.0)
    # L of welded boxes falling from height
    l0 = body(DYNAMIC, 12.0, 8.0, u(-0.3, 0.3))
    box(l0, 1.0, 0.2)
    l1 = body(DYNAMIC, 13.2, 8.8, 0.0)
    box(l1, 0.2, 1.0)
    l2 = body(DYNAMIC, 10.8, 8.8, 0.0)
    box(l2, 0.2, 0.6)
    joint(WELD, l0, l1, local_anchor_a=(1.0, 0.0), local_anchor_b=(-0.2, -0.8), reference_angle=0.0)
    joint(WELD, l0, l2, local_anchor_a=(-1.0, 0.0), local_anchor_b=(0.2, -0.8), frequency_hz=5.0,
          damping_ratio=0.3)
    # elevator on a vertical prismatic joint at x = 18
    el = body(DYNAMIC, 18.0, 2.0)
    box(el, 1.0, 0.2, density=2.0)
    joint(PRISMATIC, g, el, local_anchor_a=(18.0, 2.0), local_axis_a=(0.0, 2.0), enable_limit=1, lower=-1.5,
          upper=4.0, enable_motor=1, motor_speed=u(0.5, 2.0), max_motor=500.0)
    # pulley between two boxes
    pa = body(DYNAMIC, 22.0, 5.0)
    box(pa, 0.5, 0.5, density=u(1.0, 3.0))
    pb = body(DYNAMIC, 25.0, 6.0)
    box(pb, 0.5, 0.5, density=1.0)
    joint(PULLEY, pa, pb, ground_anchor_a=(22.0, 10.0), ground_anchor_b=(25.0, 10.0), length_a=5.0, length_b=4.0,
          ratio=u(0.8, 1.5))
    # box sliding on the ground against a friction joint
    fr = body(DYNAMIC, 28.0, 0.51, 0.0, u(3.0, 6.0))
    box(fr, 0.5, 0.5)
    joint(FRICTION, g, fr, local_anchor_a=(28.0, 0.5), max_force=3.0, max_torque=1.0, collide_connected=1)
    # box driven to an offset by a motor joint
    mo = body(DYNAMIC, 32.0, 3.0)
    box(mo, 0.4, 0.4)
    joint(MOTOR, g, mo, linear_offset=(33.0, 5.0), angular_offset=0.7, max_force=100.0, max_torque=30.0,
          correction_factor=0.3)
    # box held by a mouse joint
    ms = body(DYNAMIC, 36.0, 4.0, 0.3)
    box(ms, 0.5, 0.3)
    j = joint_defs(1, MOUSE)[0]
    j["body_a"], j["body_b"] = g, ms
    j["target"] = (36.4, 4.2)
    j["max_force"] = 500.0
    joints.append(j)
    # hinged bar with a limit and a motor
    hb = body(DYNAMIC, -36.0, 5.0)
    box(hb, 1.0, 0.1)
    joint(REVOLUTE, g, hb, local_anchor_a=(-37.0, 5.0), local_anchor_b=(-1.0, 0.0), enable_limit=1, lower=-0.7,
          upper=0.3, enable_motor=1, motor_speed=1.0, max_motor=5.0)
    b = np.array(bodies, BODY_DTYPE)
    fx = np.array(fixtures, FIXTURE_DTYPE)
    return Scene(b, fx, name=f"joints{seed}", joints=np.array(joints, JOINT_DTYPE))
