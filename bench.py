#!/usr/bin/env python3
"""Throughput of the batched b2World.Step hot path on one MI355X.

Workload: 4096 independent worlds (lognormal body counts, median 64, sigma 0.8,
clipped to [8, 1024]) of convex polygons and circles dropped into 40 m x 80 m
U containers, stepped together at dt = 1/60 with 8 velocity and 3 position
iterations, with dbox's defaults: warm starting, sleeping and continuous physics
(b2World.SolveTOI) on, on the GPU and in the CPU baseline alike.  The same steps
without continuous physics are timed once more and reported beside it.

Prints one JSON line.  value = dynamic bodies (awake or asleep) x steps / wall
seconds over steps 1..K of freshly built worlds (after a separate warm-up run
that also picks the workgroup sizes), with the worlds already resident in HBM.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from dbox_amd import native, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {"collide": "k_collide", "islands": "k_islands", "solve": "k_solve_wave", "broadphase": "k_broadphase"}
DT = 1.0 / 60.0
VEL, POS = 8, 3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def shape_bytes(fixtures: np.ndarray) -> int:
    """Shape bytes read per proxy: 12 + 16 per polygon vertex, 20 per circle, 36 per edge."""
    kind = fixtures["shape"]
    nv = np.where(kind == scenes.SHAPE_BOX, 4, fixtures["count"])
    b = np.where(kind == scenes.SHAPE_CIRCLE, 20, np.where(kind == scenes.SHAPE_EDGE, 36, 12 + 16 * nv))
    return int(b.sum())


def algorithmic_bytes(kernel: str, t: dict, nbodies: int, ndyn: int, nprox: int, sbytes: int, npairs: int) -> float:
    """Compulsory HBM bytes of one launch of each step kernel (per step)."""
    c, touch = t["contacts"], t["touching"]
    if kernel == "solve":
        # bodies: read c,a,v,w,F,tau,invMass,invI,localCenter,gravityScale,dampings,flags/sleep,fat AABB (92 B),
        # write c,a,v,w,xf,c0,a0,flags/sleep,fat AABB (76 B); touching contacts: read manifold (84 B), write impulses (16 B)
        return ndyn * (92 + 76) + touch * (84 + 16) + sbytes
    if kernel == "collide":
        # every contact read and written once (84 B each way), shapes and body transforms read once
        return c * (84 + 84) + sbytes + nbodies * 16
    if kernel == "broadphase":
        # fat AABBs and moved flags read, unique pairs written
        return nprox * 20 + npairs * 8
    # islands: contact keys and flags, body flags
    return c * 12 + nbodies * 8


def run_cpu_baseline(worlds, steps, threads, sample_every, continuous):
    """Steps 1..steps of every sample_every-th world on the oracle (C++ restatement
    of dbox), single-threaded and on a pool of threads, one world per task."""
    from oracle import pyoracle as po
    sample = worlds[::sample_every]
    dyn = sum(s.dynamic_count for s in sample)

    def run(ws, n):
        for _ in range(n):
            for w in ws:
                w.step(DT, VEL, POS)

    single = [po.OracleWorld(s, continuous=continuous) for s in sample]
    t0 = time.perf_counter()
    run(single, steps)
    t1 = time.perf_counter() - t0

    ows = [po.OracleWorld(s, continuous=continuous) for s in sample]
    parts = [ows[i::threads] for i in range(threads)]
    ts = [threading.Thread(target=run, args=(p, steps)) for p in parts]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    tn = time.perf_counter() - t0
    return dyn, t1, tn, len(sample)


def learned_vs_plain(n_eval: int, steps: int) -> dict:
    """Velocity iterations to a 1e-3 m/s residual (cap 100), plain warm start vs the
    learned one, on the first n_eval worlds of the 16,384-world stack set (tall
    uneven stacks, unseen in training).  The network is trained here, on the GPU,
    on impulses logged from plain-solver runs of other seeds (pyramids, stacks and
    containers iterated to a 1e-5 residual)."""
    from dbox_amd import warmstart
    t0 = time.perf_counter()
    h = warmstart.collect(warmstart.training_scenes(seed=11, pyramids=4, stacks=48, mixed=16), steps=150)
    samples = h.mlp_log_count()
    params, norm, loss = warmstart.train(h, steps=2000, batch=4096, lr=2e-3)
    h.close()
    test = scenes.stack_worlds(n_eval)
    plain = warmstart.iterations(test, steps)
    learned = warmstart.iterations(test, steps, params=params, norm=norm)

    def row(st):
        return {"mean_iterations": st.mean_iterations, "p95_iterations": st.p95_iterations,
                "capped_fraction": st.capped_fraction, "max_penetration_m": st.max_penetration,
                "world_steps": st.world_steps}

    return {
        "workload": f"first {n_eval} worlds of the 16,384 stack-world set, steps 1-{steps}",
        "residual_tolerance": 1e-3, "max_iterations": 100,
        "plain": row(plain), "learned": row(learned),
        "iteration_reduction": 1.0 - learned.mean_iterations / max(plain.mean_iterations, 1e-9),
        "training": {"samples": int(samples), "adam_steps": 2000, "batch": 4096,
                     "loss_first_20": float(np.mean(loss[:20])), "loss_last_20": float(np.mean(loss[-20:]))},
        "seconds": time.perf_counter() - t0,
    }


def single_world(which: int, steps: int, warmup: int, no_cpu: bool) -> dict:
    """One big world on its own: config 2 (20,000 bodies in a 200 m x 400 m
    container) or config 4 (the 50,000-box pyramid), dbox's defaults (warm
    starting, sleeping, continuous physics), exact order.  Steps warmup+1 ..
    warmup+steps are timed on the GPU with the per-phase split, and the same
    steps on the oracle (one thread) as the CPU baseline."""
    sc = scenes.big_mixed_world() if which == 2 else scenes.big_pyramid()
    flags = native.WARM | native.SLEEP | native.ORDERED | native.CCD
    h = native.StepWorlds()
    h.add_scene(sc)
    h.finalize()
    if warmup:
        h.step(DT, VEL, POS, flags, n=warmup)
    h.sync()
    h.profile_totals(reset=True)
    t0 = time.perf_counter()
    h.step(DT, VEL, POS, flags, n=steps)
    h.sync()
    t = time.perf_counter() - t0
    prof, k = h.profile_totals(reset=True)
    ndyn = sc.dynamic_count
    out = {
        "metric": f"body-steps/s, single world (config {which})", "value": ndyn * steps / t, "unit": "body-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * t / steps, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": ("20,000 bodies dropped into a 200 m x 400 m U container" if which == 2 else
                                "pyramid of 50,000 1 m boxes on a ground edge"),
                   "dynamic_bodies": ndyn, "continuous": True, "solver": "ordered"},
        "phase_ms_per_step": {kk: v / max(k, 1) for kk, v in prof.items()},
        "toi_events_last_step": int(h.toi_events()[0]),
    }
    h.close()
    if not no_cpu:
        from oracle import pyoracle as po
        o = po.OracleWorld(sc, continuous=True)
        if warmup:
            o.step(DT, VEL, POS, n=warmup)
        t0 = time.perf_counter()
        o.step(DT, VEL, POS, n=steps)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": ndyn * steps / tc, "unit": "body-steps/s", "cores": 1, "kind": "port",
                               "ms_per_step": 1e3 * tc / steps,
                               "sample": f"the same world and steps {warmup + 1}-{warmup + steps} on the C++ "
                                         "restatement of dbox (-O2 -ffp-contract=off), one thread"}
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600,
                    help="timed steps, from step 1 of freshly built worlds")
    ap.add_argument("--warmup", type=int, default=30,
                    help="steps of the separate warm-up run (which also autotunes the workgroup sizes)")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--coloured", action="store_true",
                    help="time the greedy-coloured solver instead of the reference's exact Gauss-Seidel order")
    ap.add_argument("--cpu-sample-every", type=int, default=32)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-vs-plain warm-start comparison (iterations to a fixed residual)")
    ap.add_argument("--no-ccd", action="store_true",
                    help="continuous physics off on both sides (dbox's default is on)")
    ap.add_argument("--no-discrete", action="store_true",
                    help="skip the extra timed run without continuous physics")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the second timed run whose steps are replayed from captured hipGraphs")
    ap.add_argument("--subset", default="", help="largest:N or smallest:N of the generated worlds (timing studies)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run 20 device-to-device copies of this many MiB (HBM counter calibration) and exit")
    ap.add_argument("--single-world", type=int, choices=(2, 4), default=0,
                    help="time one big world instead (config 2: 20,000-body container; config 4: 50,000-box "
                         "pyramid) with --steps timed steps after --warmup, beside the oracle on one thread")
    args = ap.parse_args()
    if args.single_world:
        if args.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            sys.exit(2)
        print(json.dumps(single_world(args.single_world, args.steps, args.warmup, args.no_cpu)))
        return
    if args.copy_mib:
        import torch
        x = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(20):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "copies": 20}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (worlds shard across GPUs with no "
              "exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    t_gen = time.perf_counter()
    worlds = scenes.batched_worlds(args.worlds)
    log(f"generated {len(worlds)} worlds in {time.perf_counter() - t_gen:.1f}s")
    if args.subset:
        kind, n = args.subset.split(":")
        by_size = sorted(range(len(worlds)), key=lambda i: len(worlds[i].bodies), reverse=(kind == "largest"))
        worlds = [worlds[i] for i in sorted(by_size[: int(n)])]
        log(f"subset {args.subset}: bodies per world {[len(w.bodies) for w in worlds][:8]}...")
    def build():
        handle = native.StepWorlds()
        for sc in worlds:
            handle.add_scene(sc)
        handle.finalize()
        return handle

    ndyn = sum(s.dynamic_count for s in worlds)
    nbodies = sum(len(s.bodies) for s in worlds)
    nprox = sum(len(s.fixtures) for s in worlds)
    sbytes = sum(shape_bytes(s.fixtures) for s in worlds)
    flags = native.WARM | native.SLEEP | (0 if args.coloured else native.ORDERED) | (0 if args.no_ccd else native.CCD)
    log(f"bodies {nbodies} dynamic {ndyn} proxies {nprox}")

    try:
        import torch
        have_torch = torch.cuda.is_available()
    except Exception:  # torch is only used for an extra device barrier
        have_torch = False

    # Warm-up run on its own copy of the worlds; its first 5 x 2 steps let the
    # autotuner pick each tier's workgroup size, which the timed run reuses.
    h = build()
    if not args.no_autotune and args.warmup >= 10:
        h.autotune(2)
    if args.warmup > 0:
        h.step(DT, VEL, POS, flags, n=args.warmup)
    h.sync()
    blocks = [int(r[3]) for r in h.tiers()[0]]
    h.close()

    # Timed run: steps 1..K of freshly built worlds.
    h = build()
    h.set_tier_blocks(blocks)
    h.sync()
    if have_torch:
        torch.cuda.synchronize()
    h.profile_totals(reset=True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        h.step(DT, VEL, POS, flags)
        if (i + 1) % 100 == 0:
            log(f"step {i + 1}/{args.steps}")
    h.sync()
    if have_torch:
        torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof, nsteps = h.profile_totals(reset=True)
    tot = h.totals()
    value = ndyn * args.steps / elapsed

    # Per-phase times are those of the lead tier (the largest worlds, launched
    # first on the handle's own stream); every other tier runs concurrently.
    kern_ms = {"collide": prof["collide"], "islands": prof["solveInit"], "solve": prof["solveVelocity"],
               "broadphase": prof["broadphase"]}
    dom = max(kern_ms, key=kern_ms.get)
    avg_ms = kern_ms[dom] / max(nsteps, 1)
    tier_rows, order = h.tiers()
    lead = [int(w) for w in order[tier_rows[0][0]:tier_rows[0][0] + tier_rows[0][1]]]
    lead_t = {"contacts": 0, "touching": 0}
    for w in lead:
        info = h.info(w)
        lead_t["contacts"] += int(info[2])
        lead_t["touching"] += int(info[3])
    abytes = algorithmic_bytes(dom, lead_t, sum(len(worlds[w].bodies) for w in lead),
                               sum(worlds[w].dynamic_count for w in lead), sum(len(worlds[w].fixtures) for w in lead),
                               sum(shape_bytes(worlds[w].fixtures) for w in lead),
                               sum(h.pairs(w).shape[0] for w in lead))
    achieved = abytes / (avg_ms * 1e-3) / 1e9
    npairs = sum(h.pairs(w).shape[0] for w in range(len(worlds)))
    step_bytes = sum(algorithmic_bytes(k, tot, nbodies, ndyn, nprox, sbytes, npairs) for k in kern_ms)
    log("lead-tier ms/step: " + ", ".join(f"{k} {v / max(nsteps, 1):.3f}" for k, v in kern_ms.items()) +
        f"; tiers {tier_rows.tolist()}")

    # HBM bytes per launch of the lead-tier solver from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # summarised in profiles/ (counters cannot be read from inside the timed run)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_solve_wave_lead_4096worlds_traffic.json")
    if dom == "solve" and not args.coloured and not args.subset and len(worlds) == 4096 and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes"]
    toi_ms = prof["solveTOI"] / max(nsteps, 1)
    toi_events = int(h.toi_events().sum())
    graph = None
    if not args.no_graph:
        # the same steps 1..K once more on fresh worlds, each tier's step replayed
        # from a captured hipGraph (captured during steps 1 and 2)
        h.close()
        h = build()
        h.set_tier_blocks(blocks)
        h.set_graph(True)
        h.sync()
        if have_torch:
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.step(DT, VEL, POS, flags, n=args.steps)
        h.sync()
        if have_torch:
            torch.cuda.synchronize()
        tg = time.perf_counter() - t0
        graph = {"value": ndyn * args.steps / tg, "ms_per_step": 1e3 * tg / args.steps,
                 "steps": f"1-{args.steps}, capture included"}
    discrete = None
    if not args.no_ccd and not args.no_discrete:
        # the same steps 1..K on fresh worlds with continuous physics off
        h.close()
        h = build()
        h.set_tier_blocks(blocks)
        h.sync()
        if have_torch:
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.step(DT, VEL, POS, flags & ~native.CCD, n=args.steps)
        h.sync()
        if have_torch:
            torch.cuda.synchronize()
        td = time.perf_counter() - t0
        discrete = {"value": ndyn * args.steps / td, "ms_per_step": 1e3 * td / args.steps,
                    "steps": f"1-{args.steps}, continuous physics off"}
    result = {
        "metric": "body-steps/s at 4096 worlds; velocity iters to fixed residual, learned vs plain",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{len(worlds)} independent 40x80 m container worlds, lognormal body counts (median 64, "
                        "sigma 0.8, clip 8-1024), 60% convex polygons 3-8 vertices / 40% circles",
            "worlds": len(worlds),
            "bodies": nbodies,
            "dynamic_bodies": ndyn,
            "dt": DT,
            "velocity_iterations": VEL,
            "position_iterations": POS,
            "warm_starting": True,
            "sleep": True,
            "continuous": not args.no_ccd,
            "solver": "coloured" if args.coloured else "ordered",
            "contacts_end": tot["contacts"],
            "touching_end": tot["touching"],
            "awake_end": tot["awake"],
        },
        "roofline": {
            "kernel": ("k_solve" if args.coloured and dom == "solve" else KERNEL_NAMES[dom]) + "<true> (lead tier: %d largest worlds)" % len(lead),
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": abytes,
            "avg_kernel_ms": avg_ms,
        },
        "lead_tier_phase_ms_per_step": {k: v / max(nsteps, 1) for k, v in kern_ms.items()},
        "tiers": [{"worlds": int(r[1]), "lds": int(r[2]), "block": int(r[3])} for r in tier_rows],
        "graph_replay": graph,
        "discrete_only": discrete,
        "continuous_pass": {"lead_tier_ms_per_step": toi_ms, "toi_events_last_step": toi_events},
        "step_algorithmic_gbs": step_bytes / (elapsed / args.steps) / 1e9,
    }
    if not args.no_learned and not args.subset:
        h.close()
        log("learned vs plain warm start")
        result["learned_vs_plain"] = learned_vs_plain(256, 150)
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        dyn, t1, tn, nsample = run_cpu_baseline(worlds, args.steps, threads, args.cpu_sample_every, not args.no_ccd)
        result["cpu_baseline"] = {
            "value": dyn * args.steps / tn,
            "unit": "body-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"every {args.cpu_sample_every}th world ({nsample} worlds, {dyn} dynamic bodies), the same "
                      f"steps 1-{args.steps} with the same flags (continuous physics "
                      f"{'off' if args.no_ccd else 'on'}), C++ restatement of dbox -O2 -ffp-contract=off, one world "
                      "per task on a thread pool",
            "single_thread_value": dyn * args.steps / t1,
        }
        result["gpu_over_cpu"] = value / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
