"""bench.py's rocprofv3 summarizers on synthetic CSV files of rocprofv3's layout: the
traffic JSON takes the mean counter value of the kernel's last dispatches and scales
it by the copy calibration; the trace table averages the last dispatches."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KN = "dbx::k_solve_wave<true>(dbx::Dev, int, dbx::StepParams, int, int)"
OTHER = "dbx::k_collide<true>(dbx::Dev, int, int, int)"


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=header)
        w.writeheader()
        w.writerows(rows)


def _pmc(d, counter, per_kernel):
    rows, did = [], 0
    for name, values in per_kernel:
        for v in values:
            did += 1
            # two rows per dispatch (e.g. per XCD dimension) that the summary adds up
            for part in (0.25, 0.75):
                rows.append({"Dispatch_Id": did, "Kernel_Name": name, "Counter_Name": counter,
                             "Counter_Value": v * part})
    _write(os.path.join(d, "host", "123_pmc_counter_collection.csv"),
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], rows)


def test_traffic_summary_scales_the_last_dispatches(tmp_path):
    f, w, cf, cw = (str(tmp_path / n) for n in ("f", "w", "cf", "cw"))
    # 5 warm-up dispatches at 9999 KiB, then 3 timed ones at 100, 200, 300 KiB
    _pmc(f, "FETCH_SIZE", [(KN, [9999] * 5 + [100, 200, 300]), (OTHER, [7] * 8)])
    _pmc(w, "WRITE_SIZE", [(KN, [9999] * 5 + [50, 50, 50])])
    # a 1 MiB copy: FETCH_SIZE reports half of its 1024 KiB, WRITE_SIZE all of it
    _pmc(cf, "FETCH_SIZE", [("fill", [0]), ("copy_kernel", [512] * 20)])
    _pmc(cw, "WRITE_SIZE", [("fill", [1024]), ("copy_kernel", [1024] * 20)])
    out = str(tmp_path / "t.json")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize-pmc", f, w, cf, cw,
                        "--kernel", "k_solve_wave<true>", "--last", "3", "--copy-mib", "1", "--out", out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    t = json.load(open(out))
    assert t["fetch_scale"] == 2.0 and t["write_scale"] == 1.0
    assert t["fetch_bytes"] == 200 * 1024 * 2 and t["write_bytes"] == 50 * 1024
    assert t["traffic_bytes"] == t["fetch_bytes"] + t["write_bytes"]
    assert any("--pmc FETCH_SIZE" in c for c in t["commands"])


def test_traffic_summary_each_matching_kernel(tmp_path):
    f, w, cf, cw = (str(tmp_path / n) for n in ("f", "w", "cf", "cw"))
    # both lead-tier kernels counted in the same runs (filtered by "<true>"), another one not matched
    _pmc(f, "FETCH_SIZE", [(KN, [9999] * 2 + [100, 300]), (OTHER, [9999] * 2 + [10, 30]),
                           ("dbx::k_toi<false>(dbx::Dev)", [5] * 4)])
    _pmc(w, "WRITE_SIZE", [(KN, [9999] * 2 + [40, 40]), (OTHER, [9999] * 2 + [8, 8]),
                           ("dbx::k_toi<false>(dbx::Dev)", [5] * 4)])
    _pmc(cf, "FETCH_SIZE", [("copy_kernel", [512] * 20)])
    _pmc(cw, "WRITE_SIZE", [("copy_kernel", [1024] * 20)])
    out = str(tmp_path / "t.json")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize-pmc", f, w, cf, cw,
                        "--kernel", "<true>", "--each", "--last", "2", "--copy-mib", "1", "--out", out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    k = json.load(open(out))["kernels"]
    assert sorted(k) == ["k_collide<true>", "k_solve_wave<true>"]
    assert k["k_solve_wave<true>"]["fetch_bytes"] == 200 * 1024 * 2
    assert k["k_solve_wave<true>"]["write_bytes"] == 40 * 1024
    assert k["k_collide<true>"]["traffic_bytes"] == 20 * 1024 * 2 + 8 * 1024


def test_trace_summary_averages_the_last_dispatches(tmp_path):
    d = str(tmp_path / "kt")
    rows, did = [], 0
    for name, durs in ((KN, [10_000] * 4 + [2_000, 4_000]), (OTHER, [500] * 6)):
        for ns in durs:
            did += 1
            rows.append({"Dispatch_Id": did, "Kernel_Name": name, "Start_Timestamp": 1000 * did,
                         "End_Timestamp": 1000 * did + ns})
    _write(os.path.join(d, "host", "1_run_kernel_trace.csv"),
           ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"], rows)
    _write(os.path.join(d, "host", "1_run_kernel_stats.csv"),
           ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"],
           [{"Name": KN, "Calls": 6, "TotalDurationNs": 46000, "AverageNs": 7666.7, "Percentage": 94.0}])
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize-trace", d, "--last", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    row = [ln for ln in r.stdout.splitlines() if ln.startswith("| `k_solve_wave<true>` | 2 ")]
    assert row and "| 3.0 | 2.0 | 4.0 |" in row[0], r.stdout


def test_step_summary_sums_each_kernel_over_the_timed_steps(tmp_path):
    # --summarize-steps: every kernel's dispatches from the one starting with the
    # (--skip + 1)-th k_collide<true> dispatch on, summed and divided by --last
    d = str(tmp_path / "kt")
    wide = "dbx::k_iw_dfs(dbx::Dev, int, int)"
    rows, t = [], 0
    for step in range(5):  # 2 untimed steps, then 3 timed ones
        for name, ns in ((OTHER, 1_000 + step), (wide, 2_000), (wide, 3_000), (KN, 10_000 * (step + 1))):
            t += 100_000
            rows.append({"Dispatch_Id": len(rows) + 1, "Kernel_Name": name, "Start_Timestamp": t,
                         "End_Timestamp": t + ns})
    _write(os.path.join(d, "host", "1_run_kernel_trace.csv"),
           ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"], rows)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize-steps", d, "--skip", "2",
                        "--last", "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = {ln.split("`")[1]: ln for ln in r.stdout.splitlines() if ln.startswith("| `")}
    # k_solve_wave<true>: (30 + 40 + 50) us / 3 steps; k_iw_dfs: 2 dispatches per step of 2 + 3 us
    assert "| 40.0 | 3 |" in lines["k_solve_wave<true>"], r.stdout
    assert "| 5.0 | 6 |" in lines["k_iw_dfs"], r.stdout
    assert "| 1.0 | 3 |" in lines["k_collide<true>"], r.stdout
    assert "over 3 collide dispatches" in r.stdout
