#!/usr/bin/env python3
"""Throughput of the batched b2World.Step hot path on one MI355X.

Workload: 4096 independent worlds (lognormal body counts, median 64, sigma 0.8,
clipped to [8, 1024]) of convex polygons and circles dropped into 40 m x 80 m
U containers, stepped together at dt = 1/60 with 8 velocity and 3 position
iterations, with dbox's defaults: warm starting, sleeping and continuous physics
(b2World.SolveTOI) on, on the GPU and in the CPU baseline alike.  The same steps
without continuous physics are timed once more and reported beside it.

Prints one JSON line.  value = dynamic bodies (awake or asleep) x steps / wall
seconds over K timed steps, with the worlds already resident in HBM, after a
separate warm-up run that also picks the workgroup sizes.  With K >= 600 the
timed steps are steps 1..K of freshly built worlds (the measurement protocol:
steps 1-600); with fewer, the worlds are first stepped untimed to step 300, the
middle of that protocol, and steps 301..300+K are timed, so a short run measures
piles rather than the first free-fall steps (--start overrides the step the timed
window starts after).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from dbox_amd import native, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {"collide": "k_collide", "islands": "k_islands", "solve": "k_solve_wave", "broadphase": "k_broadphase",
                "toi": "k_toi"}
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


def _rocprof_rows(directory: str, suffix: str) -> list[dict]:
    """Rows of every rocprofv3 CSV file under directory whose name ends in suffix."""
    import csv
    import glob
    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)):
        with open(path, newline="") as f:
            rows += list(csv.DictReader(f))
    if not rows:
        raise SystemExit(f"no *{suffix} under {directory}")
    return rows


def _per_dispatch(rows: list[dict], counter: str | None = None) -> dict:
    """{kernel name: {dispatch id: value}} (counter rows summed per dispatch; trace rows: duration in ns)."""
    out: dict = {}
    for r in rows:
        if counter is not None:
            if r["Counter_Name"] != counter:
                continue
            v = float(r["Counter_Value"])
        else:
            v = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
        d = out.setdefault(r["Kernel_Name"], {})
        k = int(r["Dispatch_Id"])
        d[k] = d.get(k, 0.0) + v
    return out


def _kernel_values(table: dict, kernel: str, last: int) -> list[float]:
    """Values of the last `last` dispatches of the kernel whose name starts with `kernel`
    (rocprofv3 writes qualified names: dbx::k_solve_wave<true>(...))."""
    hits = [n for n in table if n.split("(")[0].endswith(kernel)]
    if len(hits) != 1:
        raise SystemExit(f"kernel {kernel!r}: matches {hits}")
    d = table[hits[0]]
    return [d[k] for k in sorted(d)][-last:]


def summarize_counters(args) -> dict:
    """HBM traffic per launch of one kernel from four rocprofv3 --pmc runs (FETCH_SIZE and
    WRITE_SIZE of the profiled bench run, then of a plain device copy of --copy-mib MiB
    each), written as the JSON file bench.py reads into roofline.traffic.  With --each,
    --kernel is the regular expression the runs were filtered with, and every kernel it
    matches gets its own entry under "kernels"."""
    fetch_dir, write_dir, cfetch_dir, cwrite_dir = args.summarize_pmc
    mib = args.copy_mib or 256
    out = {"kernel": args.kernel, "workload": args.workload, "dispatches": args.last}
    tables, each = {}, {}
    for counter, d, cd in (("FETCH_SIZE", fetch_dir, cfetch_dir), ("WRITE_SIZE", write_dir, cwrite_dir)):
        table = _per_dispatch(_rocprof_rows(d, "counter_collection.csv"), counter)
        # the copy run launches one kernel exactly 20 times: the copies
        copies = [v for v in _per_dispatch(_rocprof_rows(cd, "counter_collection.csv"), counter).values()
                  if len(v) == 20]
        if len(copies) != 1:
            raise SystemExit(f"{cd}: expected one kernel with 20 dispatches")
        ckib = float(np.median(list(copies[0].values())))
        scale = mib * 1024.0 / ckib
        key = counter.split("_")[0].lower()
        out[f"{key}_copy_kib"] = ckib
        out[f"{key}_scale"] = scale
        tables[key] = (table, scale)
    if args.each:
        names = sorted({n.split("(")[0].replace("dbx::", "").replace("void ", "") for n in tables["fetch"][0]
                        if re.search(args.kernel, n.split("(")[0])})
        for name in names:
            e = {}
            for key, (table, scale) in tables.items():
                vals = _kernel_values(table, name, args.last)
                e[f"{key}_kib_mean"] = float(np.mean(vals))
                e[f"{key}_bytes"] = int(round(np.mean(vals) * 1024.0 * scale))
            e["traffic_bytes"] = e["fetch_bytes"] + e["write_bytes"]
            each[name] = e
        out["kernels"] = each
    else:
        for key, (table, scale) in tables.items():
            vals = _kernel_values(table, args.kernel, args.last)
            out[f"{key}_kib_mean"] = float(np.mean(vals))
            out[f"{key}_bytes"] = int(round(np.mean(vals) * 1024.0 * scale))
        out["traffic_bytes"] = out["fetch_bytes"] + out["write_bytes"]
    pa = args.profiled_args
    rx = args.profiled_regex or args.kernel
    out["commands"] = [
        f"rocprofv3 --pmc FETCH_SIZE --kernel-include-regex '{rx}' --output-format csv -d {fetch_dir} "
        f"-o pmc -- python3 bench.py {pa}",
        f"rocprofv3 --pmc WRITE_SIZE --kernel-include-regex '{rx}' --output-format csv -d {write_dir} "
        f"-o pmc -- python3 bench.py {pa}",
        f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {cfetch_dir} -o copy -- python3 bench.py --copy-mib {mib}",
        f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {cwrite_dir} -o copy -- python3 bench.py --copy-mib {mib}",
        "python3 bench.py --summarize-pmc " + " ".join(args.summarize_pmc) +
        f" --kernel '{args.kernel}'" + (" --each" if args.each else "") +
        (f" --profiled-regex '{args.profiled_regex}'" if args.profiled_regex else "") +
        f" --last {args.last} --copy-mib {mib} --workload '{args.workload}' --profiled-args '{pa}' --out {args.out}",
    ]
    out["note"] = ("scale = bytes the copy moves each way / what the counter reports for it; "
                   "bytes = mean counter value of the last dispatches (the timed window) x 1024 x scale")
    return out


def summarize_trace(args) -> str:
    """Markdown table of the lead tier's kernels from a rocprofv3 --kernel-trace run: mean
    duration over the last --last dispatches of each, plus rocprofv3's own stats rows."""
    table = _per_dispatch(_rocprof_rows(args.summarize_trace, "kernel_trace.csv"))
    lines = [f"| kernel | dispatches averaged | mean µs | min µs | max µs |", "|---|---|---|---|---|"]
    for name in sorted(table, key=lambda n: -sum(table[n].values())):
        short = name.split("(")[0].replace("dbx::", "")
        if not short.endswith("<true>"):
            continue
        vals = [table[name][k] for k in sorted(table[name])][-args.last:]
        lines.append(f"| `{short}` | {len(vals)} | {np.mean(vals) / 1e3:.1f} | {np.min(vals) / 1e3:.1f} | "
                     f"{np.max(vals) / 1e3:.1f} |")
    stats = _rocprof_rows(args.summarize_trace, "kernel_stats.csv")
    lines += ["", "rocprofv3 `--stats` (whole run, every dispatch including warm-up and pre-roll):", "",
              "| kernel | calls | average µs | total ms | % |", "|---|---|---|---|---|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:16]:
        lines.append(f"| `{r['Name'].split('(')[0].replace('dbx::', '')}` | {r['Calls']} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['TotalDurationNs']) / 1e6:.1f} | "
                     f"{float(r['Percentage']):.1f} |")
    return "\n".join(lines)


def summarize_steps(args) -> str:
    """Markdown table of every kernel of a --single-world rocprofv3 --kernel-trace run: the
    time each takes per timed step, summed over its dispatches from the one that starts
    with the (--skip + 1)-th k_collide dispatch on (the first timed step), divided by --last."""
    rows = _rocprof_rows(args.summarize_steps, "kernel_trace.csv")

    def name(r):
        return r["Kernel_Name"].split("(")[0].replace("dbx::", "").replace("void ", "")

    starts = sorted(int(r["Start_Timestamp"]) for r in rows if name(r) == "k_collide<true>")
    t0 = starts[args.skip]
    tot, cnt = {}, {}
    for r in rows:
        if int(r["Start_Timestamp"]) >= t0:
            k = name(r)
            tot[k] = tot.get(k, 0) + int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            cnt[k] = cnt.get(k, 0) + 1
    lines = ["| kernel | µs per step (timed steps) | dispatches |", "|---|---|---|"]
    for k in sorted(tot, key=lambda k: -tot[k]):
        lines.append(f"| `{k}` | {tot[k] / args.last / 1e3:.1f} | {cnt[k]} |")
    lines += ["", f"sum of kernel time: {sum(tot.values()) / args.last / 1e3:.1f} µs per step over "
              f"{len(starts) - args.skip} collide dispatches in the window"]
    return "\n".join(lines)


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
    if kernel == "toi":
        # the pass without events: every contact's TOI record written (16 B) and read by
        # the first search (8 B), every body's sweep start reset (16 B)
        return c * 24 + nbodies * 16
    # islands: contact keys and flags, body flags
    return c * 12 + nbodies * 8


def host_cpu() -> dict:
    """The host the CPU baseline runs on: CPU model, logical CPUs, the CPUs this
    process may use (affinity, cgroup quota, OMP_NUM_THREADS), oracle compiler and flags."""
    info = {"nproc": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            info["model"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        info["model"] = ""
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = info["nproc"]
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            info["cgroup_cpus"] = None if q == "max" else float(q) / float(per)
    except (OSError, ValueError):
        info["cgroup_cpus"] = None
    info["omp_num_threads"] = os.environ.get("OMP_NUM_THREADS")
    try:
        import subprocess
        info["compiler"] = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        info["compiler"] = ""
    info["flags"] = "-std=c++17 -O3 -ffp-contract=off -fno-fast-math (oracle/Makefile)"
    return info


def cpu_threads(info: dict) -> int:
    """Every CPU this process is granted: affinity, capped by a cgroup quota and by
    OMP_NUM_THREADS, which the GPU hosts set to the CPU share of one GPU."""
    n = info["affinity"] or 1
    if info["cgroup_cpus"]:
        n = min(n, max(1, int(info["cgroup_cpus"])))
    if info["omp_num_threads"] and info["omp_num_threads"].isdigit():
        n = min(n, int(info["omp_num_threads"]))
    return max(1, n)


def run_cpu_baseline(worlds, start, steps, threads, sample_every, single_every, continuous):
    """Steps start+1..start+steps of every sample_every-th world on the oracle (C++
    restatement of dbox) on a pool of threads, one world per task, after stepping
    them untimed to `start`; and the same window of every single_every-th world on
    one thread."""
    from oracle import pyoracle as po

    def run(ws, n):
        for _ in range(n):
            for w in ws:
                w.step(DT, VEL, POS)

    def pool(ws, n):
        parts = [ws[i::threads] for i in range(threads)]
        ts = [threading.Thread(target=run, args=(p, n)) for p in parts]
        t0 = time.perf_counter()
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        return time.perf_counter() - t0

    sample = worlds[::sample_every]
    dyn = sum(s.dynamic_count for s in sample)
    ows = [po.OracleWorld(s, continuous=continuous) for s in sample]
    if start:
        pool(ows, start)
    tn = pool(ows, steps)

    one = worlds[::single_every]
    dyn1 = sum(s.dynamic_count for s in one)
    ows1 = [po.OracleWorld(s, continuous=continuous) for s in one]
    if start:
        run(ows1, start)
    t0 = time.perf_counter()
    run(ows1, steps)
    t1 = time.perf_counter() - t0
    return {"dyn": dyn, "t_pool": tn, "n": len(sample), "dyn1": dyn1, "t_one": t1, "n1": len(one)}


NETWORK = os.path.join(ROOT, "dbox_amd", "learned", "warmstart_c12.npz")


def learned_train(path: str, budget_s: float, target: int) -> dict:
    """Trains the learned warm start at its specified scale and saves it: new-point
    samples logged from plain-solver runs of configs 1-2 (pyramids and 20,000-body
    containers of their own seeds, iterated to a 1e-5 residual) until `target` samples
    or `budget_s` seconds, then Adam on the GPU with the library's own kernels."""
    from dbox_amd import warmstart
    t0 = time.perf_counter()
    h, info = warmstart.collect_spec(target=target, budget_s=budget_s, progress=log)
    params, norm, loss = warmstart.train_schedule(h)
    h.close()
    meta = {"data": info, "training": {"adam": "4000 steps at 2e-3, 2000 at 1e-3, 2000 at 5e-4", "batch": 8192,
                                       "loss_first_100": float(np.mean(loss[:100])),
                                       "loss_last_100": float(np.mean(loss[-100:]))},
            "command": f"python3 bench.py --learned-train {os.path.relpath(path, ROOT)} --learned-budget {budget_s:g} "
                       f"--learned-target {target}",
            "seconds": time.perf_counter() - t0}
    os.makedirs(os.path.dirname(path), exist_ok=True)
    warmstart.save_network(path, params, norm, meta)
    return meta


def learned_vs_plain(c5_steps: int = 600, c4_steps: int = 60) -> dict:
    """Velocity iterations to a 1e-3 m/s residual (at most 100 sweeps), the plain warm
    start against the learned one, over the same world-steps, and the largest
    penetration: config 5 (16,384 worlds of tall uneven stacks, steps 1-c5_steps) and
    config 4 (the 50,000-box pyramid, steps 1-c4_steps).  The network is the one
    trained at the specified scale (bench.py --learned-train), or, if that file is
    absent, one trained here on a small sample."""
    from dbox_amd import warmstart
    t0 = time.perf_counter()
    if os.path.exists(NETWORK):
        params, norm, meta = warmstart.load_network(NETWORK)
        source = {"network": os.path.relpath(NETWORK, ROOT), **meta}
    else:
        h = warmstart.collect(warmstart.training_scenes(seed=11, pyramids=4, stacks=48, mixed=16), steps=150)
        samples = h.mlp_log_count()
        params, norm, loss = warmstart.train(h, steps=2000, batch=4096, lr=2e-3)
        h.close()
        source = {"network": "trained here on a small sample", "samples": int(samples)}

    def row(st):
        return {"mean_iterations": st.mean_iterations, "p95_iterations": st.p95_iterations,
                "capped_fraction": st.capped_fraction, "max_penetration_m": st.max_penetration,
                "world_steps": st.world_steps, "awake_body_steps": st.awake_body_steps}

    out = {"residual_tolerance": 1e-3, "max_iterations": 100, "training": source}
    for name, worlds, steps in (("config5", scenes.stack_worlds(16384), c5_steps),
                                ("config4", [scenes.big_pyramid()], c4_steps)):
        if steps <= 0:
            continue
        log(f"learned vs plain warm start: {name}, {steps} steps")
        plain = warmstart.iterations(worlds, steps)
        learned = warmstart.iterations(worlds, steps, params=params, norm=norm)
        cmp = warmstart.compare(plain, learned)
        out[name] = {"workload": (f"16,384 stack worlds, steps 1-{steps}" if name == "config5" else
                                  f"pyramid of 50,000 boxes, steps 1-{steps}"),
                     "plain": row(plain), "learned": row(learned), "same_world_steps": cmp,
                     "iteration_reduction": cmp["iteration_reduction"]}
    out["iteration_reduction"] = out["config5"]["iteration_reduction"] if "config5" in out else None
    out["seconds"] = time.perf_counter() - t0
    return out


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
                                         "restatement of dbox (-O3 -ffp-contract=off), one thread"}
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600,
                    help="timed steps (steps 1..K of freshly built worlds for K >= 600, else steps 301..300+K)")
    ap.add_argument("--start", type=int, default=-1,
                    help="untimed steps of the timed run's worlds before the window (default 0 for >= 600 timed "
                         "steps, else 300)")
    ap.add_argument("--warmup", type=int, default=30,
                    help="steps of the separate warm-up run (which also autotunes the workgroup sizes)")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--coloured", action="store_true",
                    help="time the greedy-coloured solver instead of the reference's exact Gauss-Seidel order")
    ap.add_argument("--cpu-sample-every", type=int, default=8)
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
    ap.add_argument("--learned-c5-steps", type=int, default=600,
                    help="steps of the learned-vs-plain comparison on the 16,384 stack worlds (0: skip)")
    ap.add_argument("--learned-c4-steps", type=int, default=60,
                    help="steps of the learned-vs-plain comparison on the 50,000-box pyramid (0: skip)")
    ap.add_argument("--learned-train", metavar="PATH",
                    help="train the learned warm start at its specified scale, save it to PATH and exit")
    ap.add_argument("--learned-budget", type=float, default=420.0, help="seconds of data collection for --learned-train")
    ap.add_argument("--learned-target", type=int, default=10_000_000, help="samples for --learned-train")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="write the traffic JSON of --kernel from four rocprofv3 --pmc output folders and exit")
    ap.add_argument("--summarize-trace", metavar="TRACE_DIR",
                    help="print the lead tier's kernel durations from a rocprofv3 --kernel-trace folder and exit")
    ap.add_argument("--kernel", default="k_solve_wave<true>")
    ap.add_argument("--each", action="store_true", help="--summarize-pmc: one entry per kernel --kernel matches")
    ap.add_argument("--last", type=int, default=600, help="dispatches (timed steps) the summaries average over")
    ap.add_argument("--summarize-steps", metavar="TRACE_DIR",
                    help="print each kernel's time per timed step of a --single-world kernel-trace run and exit")
    ap.add_argument("--skip", type=int, default=0, help="untimed steps before the window (--summarize-steps)")
    ap.add_argument("--workload", default="")
    ap.add_argument("--profiled-args", default="")
    ap.add_argument("--profiled-regex", default="",
                    help="the --kernel-include-regex the counter runs used, when wider than --kernel")
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    if args.summarize_pmc:
        res = summarize_counters(args)
        text = json.dumps(res, indent=1)
        if args.out:
            with open(args.out, "w") as f:
                f.write(text + "\n")
        print(text)
        return
    if args.summarize_steps:
        print(summarize_steps(args))
        return
    if args.summarize_trace:
        print(summarize_trace(args))
        return
    if args.learned_train:
        print(json.dumps(learned_train(os.path.abspath(args.learned_train), args.learned_budget,
                                       args.learned_target)))
        return
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
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the worlds would be dealt out "
              "over the GPUs, one process and handle each, with no exchange between them)", file=sys.stderr)
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

    start = args.start if args.start >= 0 else (0 if args.steps >= 600 else 300)
    window = f"{start + 1}-{start + args.steps}"
    ndyn = sum(s.dynamic_count for s in worlds)
    nbodies = sum(len(s.bodies) for s in worlds)
    nprox = sum(len(s.fixtures) for s in worlds)
    sbytes = sum(shape_bytes(s.fixtures) for s in worlds)
    flags = native.WARM | native.SLEEP | (0 if args.coloured else native.ORDERED) | (0 if args.no_ccd else native.CCD)
    log(f"bodies {nbodies} dynamic {ndyn} proxies {nprox}; timed steps {window}")

    def barrier(h):
        # dbx_sync joins every tier's stream into the handle's stream and waits for it:
        # all of the handle's work on the device has finished when it returns
        h.sync()

    # Warm-up run on its own copy of the worlds; its first 5 x 2 steps let the
    # autotuner pick each tier's workgroup size, which the timed runs reuse.
    h = build()
    if not args.no_autotune and args.warmup >= 10:
        h.autotune(2)
    if args.warmup > 0:
        h.step(DT, VEL, POS, flags, n=args.warmup)
    h.sync()
    blocks = [int(r[3]) for r in h.tiers()[0]]
    h.close()

    def fresh(graph=False):
        """Freshly built worlds stepped untimed to the start of the window."""
        handle = build()
        handle.set_tier_blocks(blocks)
        if start:
            handle.step(DT, VEL, POS, flags, n=start)
        handle.set_graph(graph)
        barrier(handle)
        handle.profile_totals(reset=True)
        handle.phase_totals(reset=True)
        return handle

    # Timed run
    h = fresh()
    t0 = time.perf_counter()
    for i in range(args.steps):
        h.step(DT, VEL, POS, flags)
        if (i + 1) % 100 == 0:
            log(f"step {start + i + 1}")
    barrier(h)
    elapsed = time.perf_counter() - t0
    prof, nsteps = h.profile_totals(reset=True)
    kern, ksteps = h.phase_totals(reset=True)
    tot = h.totals()
    value = ndyn * args.steps / elapsed

    # Per-kernel times of the lead tier (the largest worlds, launched first on the
    # handle's own stream); every other tier runs concurrently.
    kern_ms = {k: kern[k] for k in ("collide", "islands", "solve", "broadphase")}
    dom = max(kern_ms, key=kern_ms.get)
    avg_ms = kern_ms[dom] / max(ksteps, 1)
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
    # every lead-tier kernel against the same roofline, with its HBM bytes per launch
    # where a FETCH_SIZE / WRITE_SIZE profile of this workload and window exists
    each_path = os.path.join(ROOT, "profiles", f"lead_tier_4096worlds_steps{window}_traffic.json")
    each_traffic = {}
    if not args.coloured and not args.subset and len(worlds) == 4096 and os.path.exists(each_path):
        with open(each_path) as f:
            each_traffic = json.load(f)["kernels"]
    lead_kernels = {}
    for k in kern:
        ms = kern[k] / max(ksteps, 1)
        b = algorithmic_bytes(k, lead_t, sum(len(worlds[w].bodies) for w in lead),
                              sum(worlds[w].dynamic_count for w in lead), sum(len(worlds[w].fixtures) for w in lead),
                              sum(shape_bytes(worlds[w].fixtures) for w in lead), sum(h.pairs(w).shape[0] for w in lead))
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else None
        lead_kernels[k] = {"ms": ms, "algorithmic_bytes": b, "achieved_gbs": gbs,
                           "frac": gbs / HBM_PEAK_GBS if gbs is not None else None,
                           "traffic": each_traffic.get(KERNEL_NAMES[k] + "<true>", {}).get("traffic_bytes")}
    npairs = sum(h.pairs(w).shape[0] for w in range(len(worlds)))
    step_bytes = sum(algorithmic_bytes(k, tot, nbodies, ndyn, nprox, sbytes, npairs) for k in kern_ms)
    log("lead-tier ms/step: " + ", ".join(f"{k} {v / max(ksteps, 1):.3f}" for k, v in kern.items()) +
        f"; tiers {tier_rows.tolist()}")
    # Latency model of the largest world's island solve: its dependency levels (the
    # exact-order schedule's batches) and the time per level of each phase, from the
    # phase stamps the solve kernel takes (b2Profile's solveInit / solveVelocity /
    # solvePosition of that world).
    sched = h.schedule(lead[0])
    levels = int(sched[:, 3].max()) + 1 if len(sched) else 0
    ph = {k: prof[k] / max(nsteps, 1) for k in ("solveInit", "solveVelocity", "solvePosition")}
    latency = {
        "world": lead[0], "constraints_last_step": int(len(sched)), "levels_per_sweep_last_step": levels,
        "phase_ms_per_step": ph,
        "us_per_velocity_level": 1e3 * ph["solveVelocity"] / max(levels * VEL, 1),
        "note": "a wave walks the levels of the largest world in order, one sweep at a time (two waves "
                "overlap consecutive velocity sweeps where the bodies' level spread allows, so solveVelocity "
                "counts from the warm start's end); a level costs its instruction chain (about 1,100 shader "
                "cycles per velocity level), not memory traffic",
    } if not args.coloured else None
    # HBM bytes per launch of the named kernel from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # summarised in profiles/ for this workload and window (counters cannot be read inside the timed run)
    traffic, traffic_src = None, None
    kname = KERNEL_NAMES[dom]
    tpath = os.path.join(ROOT, "profiles", f"{kname}_lead_4096worlds_steps{window}_traffic.json")
    if not args.coloured and not args.subset and len(worlds) == 4096 and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes"]
        traffic_src = os.path.relpath(tpath, ROOT)
    toi_ms = kern["toi"] / max(ksteps, 1)
    toi_events = int(h.toi_events().sum())
    graph = None
    if not args.no_graph:
        # the same window once more on fresh worlds, each tier's step replayed from a
        # captured hipGraph (captured during the first two timed steps)
        h.close()
        h = fresh(graph=True)
        t0 = time.perf_counter()
        h.step(DT, VEL, POS, flags, n=args.steps)
        barrier(h)
        tg = time.perf_counter() - t0
        graph = {"value": ndyn * args.steps / tg, "ms_per_step": 1e3 * tg / args.steps,
                 "steps": f"{window}, capture included"}
    discrete = None
    if not args.no_ccd and not args.no_discrete:
        # the same window on fresh worlds with continuous physics off in the window
        h.close()
        h = fresh()
        t0 = time.perf_counter()
        h.step(DT, VEL, POS, flags & ~native.CCD, n=args.steps)
        barrier(h)
        td = time.perf_counter() - t0
        discrete = {"value": ndyn * args.steps / td, "ms_per_step": 1e3 * td / args.steps,
                    "steps": f"{window}, continuous physics off in the timed steps"}
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
            "timed_steps": window,
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
            "kernel": ("k_solve" if args.coloured and dom == "solve" else kname) + "<true> (lead tier: %d largest worlds)" % len(lead),
            "bound": "hbm",
            "limiter": "latency: the dependency chain of the largest world" if dom in ("solve", "islands") else "latency",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes": abytes,
            "avg_kernel_ms": avg_ms,
        },
        "latency_model": latency,
        "lead_tier_kernel_ms_per_step": {k: v / max(ksteps, 1) for k, v in kern.items()},
        "lead_tier_kernels": lead_kernels,
        "lead_tier_traffic_source": os.path.relpath(each_path, ROOT) if each_traffic else None,
        "profile_ms_per_step": {k: v / max(nsteps, 1) for k, v in prof.items()},
        "tiers": [{"worlds": int(r[1]), "lds": int(r[2]), "block": int(r[3])} for r in tier_rows],
        "graph_replay": graph,
        "discrete_only": discrete,
        "continuous_pass": {"lead_tier_ms_per_step": toi_ms, "toi_events_last_step": toi_events},
        "step_algorithmic_gbs": step_bytes / (elapsed / args.steps) / 1e9,
    }
    h.close()
    if not args.no_learned and not args.subset:
        log("learned vs plain warm start")
        result["learned_vs_plain"] = learned_vs_plain(args.learned_c5_steps, args.learned_c4_steps)
    if not args.no_cpu:
        info = host_cpu()
        threads = args.cpu_threads or cpu_threads(info)
        log(f"cpu baseline: {threads} threads on {info['model']}")
        c = run_cpu_baseline(worlds, start, args.steps, threads, args.cpu_sample_every, 4 * args.cpu_sample_every,
                             not args.no_ccd)
        result["cpu_baseline"] = {
            "value": c["dyn"] * args.steps / c["t_pool"],
            "unit": "body-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"every {args.cpu_sample_every}th world ({c['n']} worlds, {c['dyn']} dynamic bodies), the same "
                      f"steps {window} with the same flags (continuous physics {'off' if args.no_ccd else 'on'}), C++ "
                      "restatement of dbox, one world per task on a pool of every CPU this process is granted",
            "single_thread_value": c["dyn1"] * args.steps / c["t_one"],
            "single_thread_sample": f"every {4 * args.cpu_sample_every}th world ({c['n1']} worlds), same steps, one thread",
            "host": info,
        }
        result["gpu_over_cpu"] = value / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
