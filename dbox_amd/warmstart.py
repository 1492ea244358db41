"""Learned warm start: collecting training data, training and evaluating the network.

No counterpart in the reference.  b2Contact.Update (b2contact.d:306-324) gives a
manifold point without a match in the previous manifold a zero impulse; the
native library can replace that zero by the prediction of a small network
(14 features -> 64 -> 64 -> 2, bf16 matrix cores with fp32 accumulation).  The
training data are the features of new points and the impulses the plain solver
converges to for them, logged by the solver kernels themselves (DBX_STEP_LOG)
while they iterate to a tight residual; the network is trained by the library's
own forward/backward kernels with Adam.  Everything here runs on the GPU; this
module only drives it.
"""
from __future__ import annotations

import dataclasses
import json
import time

import numpy as np

from . import native, scenes

DT = 1.0 / 60.0
VEL, POS = 8, 3
PLAIN = native.WARM | native.SLEEP | native.ORDERED | native.RESIDUAL


def _handle(scene_list) -> native.StepWorlds:
    h = native.StepWorlds()
    for s in scene_list:
        h.add_scene(s)
    h.finalize()
    return h


def collect(scene_list, steps: int, capacity: int = 1 << 22, tolerance: float = 1e-5,
            max_iterations: int = 500) -> native.StepWorlds:
    """Steps the scenes with the plain solver iterating to `tolerance` and logs
    every new point's features and converged impulses.  Returns the handle,
    whose log holds the samples."""
    h = _handle(scene_list)
    h.mlp_log_enable(capacity)
    h.set_residual(tolerance, max_iterations)
    h.step(DT, VEL, POS, PLAIN | native.LOG, n=steps)
    h.sync()
    return h


def train(h: native.StepWorlds, steps: int = 2000, batch: int = 4096, lr: float = 2e-3, seed: int = 1,
          init_seed: int = 7):
    """Fits the normalisation to the log, initialises and trains the network.
    Returns (params, norm, per-step loss)."""
    h.mlp_init(init_seed)
    h.mlp_fit_normalization()
    loss = h.mlp_train(steps, batch, lr, seed)
    params, norm = h.mlp_get()
    return params, norm, loss


@dataclasses.dataclass
class IterationStats:
    mean_iterations: float      # over (world, step) pairs that solved at least one constraint
    p95_iterations: float
    capped_fraction: float      # share of those pairs that hit max_iterations
    max_penetration: float      # over all worlds and steps
    world_steps: int
    mean_residual: float        # velocity residual after the last iteration, mean over those pairs
    per_step: np.ndarray = dataclasses.field(repr=False, default=None)  # [step, world] iterations, 0 = no solve
    awake_body_steps: int = 0   # sum over steps of the awake dynamic bodies after the step


def iterations(scene_list, steps: int, params=None, norm=None, tolerance: float = 1e-3,
               max_iterations: int = 100) -> IterationStats:
    """Velocity iterations needed to reach `tolerance`, per world and step, with
    the plain warm start (params None) or the learned one."""
    h = _handle(scene_list)
    h.set_residual(tolerance, max_iterations)
    flags = PLAIN
    if params is not None:
        h.mlp_set(params, norm)
        flags |= native.LEARNED
    its, res, pen, awake = [], [], 0.0, 0
    for _ in range(steps):
        h.step(DT, VEL, POS, flags)
        h.sync()
        it, p = h.solver_stats()
        r = h.residuals()
        its.append(it.copy())
        res.append(r[it > 0])
        pen = max(pen, float(p.max()))
        awake += h.totals()["awake"]
    h.close()
    per = np.stack(its) if its else np.zeros((0, len(scene_list)), np.int32)
    a = per[per > 0]
    if len(a) == 0:
        return IterationStats(0.0, 0.0, 0.0, pen, 0, 0.0, per, awake)
    return IterationStats(float(a.mean()), float(np.percentile(a, 95)), float(np.mean(a >= max_iterations)), pen,
                          int(len(a)), float(np.concatenate(res).mean()), per, awake)


def compare(plain: IterationStats, learned: IterationStats) -> dict:
    """Learned against plain over the same (world, step) pairs: those where both runs
    solved at least one constraint (a run that keeps a world awake where the other let
    it sleep would otherwise be averaged over a different set).  Also the world-steps
    and awake body-steps of each run, which show any change in sleeping."""
    both = (plain.per_step > 0) & (learned.per_step > 0)
    p = plain.per_step[both].astype(np.float64)
    q = learned.per_step[both].astype(np.float64)
    out = {"common_world_steps": int(both.sum()),
           "plain_mean_iterations": float(p.mean()) if p.size else 0.0,
           "learned_mean_iterations": float(q.mean()) if q.size else 0.0,
           "plain_world_steps": plain.world_steps, "learned_world_steps": learned.world_steps,
           "plain_awake_body_steps": plain.awake_body_steps, "learned_awake_body_steps": learned.awake_body_steps,
           "plain_max_penetration_m": plain.max_penetration, "learned_max_penetration_m": learned.max_penetration}
    out["iteration_reduction"] = 1.0 - out["learned_mean_iterations"] / out["plain_mean_iterations"] if p.size else 0.0
    return out


def training_scenes(seed: int = 11, pyramids: int = 8, stacks: int = 48, mixed: int = 8):
    """Plain-solver data sources: seeded pyramids (the 20-row benchmark scene),
    stack worlds and small mixed containers."""
    out = [scenes.pyramid(20, seed=seed * 1000 + i) for i in range(pyramids)]
    out += [scenes.stacks_world(seed * 7919 + i) for i in range(stacks)]
    out += [scenes.mixed_container(200, 40.0, 80.0, seed * 104729 + i) for i in range(mixed)]
    return out


def spec_sources(pyramids: int = 256, containers: int = 64, seed: int = 7000):
    """The training worlds at the learned warm start's specified scale: seeded config-1
    pyramids (20 rows, 210 boxes) and config-2 containers (20,000 bodies in a 200 m x
    400 m U container), with seeds that no evaluation scene uses."""
    out = [scenes.pyramid(20, seed=seed + i) for i in range(pyramids)]
    out += [scenes.big_mixed_world(seed=seed + 500 + i) for i in range(containers)]
    return out


def collect_spec(target: int = 10_000_000, budget_s: float = 420.0, pyramids: int = 256, containers: int = 64,
                 seed: int = 7000, tolerance: float = 1e-5, max_iterations: int = 100, progress=None):
    """Plain-solver runs of configs 1-2 iterating to `tolerance` (at most
    `max_iterations` sweeps), logging every new point's features and the impulses the
    solver reached for it, until `target` samples or `budget_s` seconds.  Steps 1-120
    are logged (the pyramids' contacts form then), steps 121-300 run with the plain
    iteration counts (the containers' bodies are still falling and create few
    contacts), then every step is logged.  Returns (handle, info)."""
    t0 = time.perf_counter()
    h = _handle(spec_sources(pyramids, containers, seed))
    h.mlp_log_enable(target + (target >> 3))
    h.set_residual(tolerance, max_iterations)
    step, logged_steps = 0, 0
    while True:
        logging = step < 120 or step >= 300
        flags = (PLAIN | native.LOG) if logging else (native.WARM | native.SLEEP | native.ORDERED)
        h.step(DT, VEL, POS, flags)
        step += 1
        logged_steps += logging
        if step % 10 == 0:
            h.sync()
            n = h.mlp_log_count()
            el = time.perf_counter() - t0
            if progress is not None and step % 50 == 0:
                progress(f"learned-warm-start data: step {step}, {n} samples, {el:.0f} s")
            if n >= target or el >= budget_s:
                break
    n = min(h.mlp_log_count(), target)
    info = {"samples": int(n), "steps": step, "logged_steps": logged_steps, "seconds": time.perf_counter() - t0,
            "sources": f"{pyramids} config-1 pyramids (seeds {seed}..{seed + pyramids - 1}) and {containers} config-2 "
                       f"containers (seeds {seed + 500}..{seed + 499 + containers}), one batch",
            "labels": f"impulses after iterating to a {tolerance:g} m/s residual, at most {max_iterations} sweeps"}
    return h, info


def train_schedule(h: native.StepWorlds, schedule=((4000, 2e-3), (2000, 1e-3), (2000, 5e-4)), batch: int = 8192,
                   init_seed: int = 7, seed: int = 1):
    """Normalisation fitted to the whole log, then Adam in stages of (steps, learning
    rate).  Returns (params, norm, per-step loss)."""
    h.mlp_init(init_seed)
    h.mlp_fit_normalization()
    losses = [h.mlp_train(n, batch, lr, seed + i) for i, (n, lr) in enumerate(schedule)]
    params, norm = h.mlp_get()
    return params, norm, np.concatenate(losses)


def save_network(path: str, params: np.ndarray, norm: np.ndarray, meta: dict) -> None:
    """The trained network, its feature normalisation and how it was made (JSON), as
    plain arrays (loadable without pickle)."""
    np.savez(path, params=np.asarray(params, np.float32), norm=np.asarray(norm, np.float32),
             meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))


def load_network(path: str):
    with np.load(path) as z:
        return z["params"], z["norm"], json.loads(bytes(z["meta"]).decode())
