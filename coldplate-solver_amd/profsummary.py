"""Summarise rocprofv3 output of bench.py into a short text profile.

Reads three rocprofv3 runs of the same bench workload, each a directory
holding either the rocpd database (*_results.db, the default output) or the
CSV files (--output-format csv):
  * --kernel-trace --stats   -> <stats>
  * --pmc FETCH_SIZE         -> <fetch>
  * --pmc WRITE_SIZE         -> <write>
and the bench's own JSON line from the kernel-trace run's log (its live,
event-timed average launch duration of the same kernel), plus, with --bench,
the JSON line of the same command run without the profiler.

The byte counters are scaled against a device-to-device copy of known size
run in the same process (bench.py --calibrate-copy MIB): the copy reads and
writes MIB MiB, so factor = MIB * 1024 / counter KiB for each counter.

Writes <out>.md (human summary with the commands) and <out>.json (the
traffic figure bench.py reports as roofline.traffic).

    python coldplate-solver_amd/profsummary.py --stats D1 --fetch D2 --write D3 \
        --log D1.log --copy-mib 256 --kernel k_pcg_q --out profiles/k_pcg_q_network_1024x512_b8
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sqlite3
import sys
from collections import defaultdict


def _one(pattern: str) -> str:
    files = sorted(glob.glob(pattern))
    if not files:
        raise SystemExit(f"no file matches {pattern}")
    return files[0]


def _db(d: str) -> str | None:
    files = sorted(glob.glob(os.path.join(d, "*_results.db")))
    return files[0] if files else None


def kernel_stats(d: str) -> list[dict]:
    db = _db(d)
    if db:
        c = sqlite3.connect(db)
        rows = c.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                         "from kernels group by name").fetchall()
        total = sum(r[2] for r in rows) or 1
        return [{"Name": r[0], "Calls": str(r[1]), "TotalDurationNs": str(r[2]), "AverageNs": str(r[3]),
                 "MinNs": str(r[4]), "MaxNs": str(r[5]), "Percentage": str(100.0 * r[2] / total)} for r in rows]
    with open(_one(os.path.join(d, "*kernel_stats.csv"))) as f:
        return list(csv.DictReader(f))


def burst_durations(d: str, sub: str, run: int) -> tuple[str, list[float]]:
    """Durations (ns) of the launches bench.py times: the first `run`
    dispatches of every maximal run of at least `run` back-to-back dispatches
    (in start order) of the kernel matching `sub`. In a SIMPLE step the
    kernel is always followed by another kernel, and a timing burst holds the
    GPU alone, so only bursts make such runs."""
    db = _db(d)
    if not db:
        raise SystemExit("burst timing needs the rocpd database of the kernel-trace run")
    c = sqlite3.connect(db)
    rows = c.execute("select name, duration from kernels order by start").fetchall()
    names = {r[0] for r in rows if sub in r[0]}
    if not names:
        raise SystemExit(f"no kernel matches {sub}")
    counts = defaultdict(int)
    for r in rows:
        if r[0] in names:
            counts[r[0]] += 1
    name = max(names, key=lambda n: counts[n])
    out, cur = [], []
    for r in rows + [("", 0)]:
        if r[0] == name:
            cur.append(float(r[1]))
            continue
        if len(cur) >= run:
            out += cur[:run]
        cur = []
    return name, out


def counter_per_dispatch(d: str, counter: str) -> dict[str, list[float]]:
    """Kernel name -> list of per-dispatch counter values (summed over the
    dimensions rocprofv3 reports separately)."""
    per = defaultdict(float)
    names = {}
    db = _db(d)
    if db:
        c = sqlite3.connect(db)
        for disp, name, val in c.execute("select dispatch_id, kernel_name, value from counters_collection "
                                         "where counter_name = ?", (counter,)):
            per[disp] += float(val)
            names[disp] = name
    else:
        with open(_one(os.path.join(d, "*counter_collection.csv"))) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                key = r["Dispatch_Id"]
                per[key] += float(r["Counter_Value"])
                names[key] = r["Kernel_Name"]
    out = defaultdict(list)
    for key, v in per.items():
        out[names[key]].append(v)
    return out


def pick(d: dict[str, list[float]], sub: str) -> tuple[str, list[float]]:
    """The kernel matching `sub` with the most dispatches (the tiling the
    run used; an autotuning pass also launches the other instantiations)."""
    hits = [(k, v) for k, v in d.items() if sub in k]
    if not hits:
        raise SystemExit(f"no kernel matching {sub}")
    return max(hits, key=lambda kv: len(kv[1]))


def bench_line(log: str) -> dict | None:
    if not log or not os.path.exists(log):
        return None
    for line in reversed(open(log, errors="replace").read().splitlines()):
        line = line.strip()
        if line.startswith("{") and '"metric"' in line:
            return json.loads(line)
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch", default="", help="FETCH_SIZE run (with --write; both or neither)")
    ap.add_argument("--write", default="")
    ap.add_argument("--log", default="")
    ap.add_argument("--bench", default="", help="JSON line of the same command run without the profiler")
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--kernel", default="k_pcg_q")
    ap.add_argument("--title", default="", help="heading of the summary (default: the kernel)")
    ap.add_argument("--workload", default="network_1024x512, 8 geometries, mixed precision")
    ap.add_argument("--commands", default="", help="text file holding the exact commands that were run")
    ap.add_argument("--burst", type=int, default=0,
                    help="compare the live timing with the rocprof durations of the timing bursts "
                         "(runs of this many back-to-back launches) instead of every launch")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()

    stats = kernel_stats(a.stats)
    counters = bool(a.fetch and a.write)
    if counters:
        if a.copy_mib <= 0:
            raise SystemExit("--copy-mib is needed with the counter runs")
        fetch = counter_per_dispatch(a.fetch, "FETCH_SIZE")
        write = counter_per_dispatch(a.write, "WRITE_SIZE")
        copy_kib = a.copy_mib * 1024.0
        # the calibration copy is the largest copyBuffer dispatch
        _, cf = pick(fetch, "copyBuffer")
        _, cw = pick(write, "copyBuffer")
        f_copy, w_copy = max(cf), max(cw)
        f_factor, w_factor = copy_kib / f_copy, copy_kib / w_copy
        kname, kf = pick(fetch, a.kernel)
        _, kw = pick(write, a.kernel)
        f_avg = sum(kf) / len(kf)
        w_avg = sum(kw) / len(kw)
        traffic = (f_avg * f_factor + w_avg * w_factor) * 1024.0

    b = bench_line(a.log)
    bu = bench_line(a.bench)
    krow = max((r for r in stats if a.kernel in r["Name"]), key=lambda r: int(r["Calls"]))
    rocprof_avg_us = float(krow["AverageNs"]) / 1e3
    n_avg = int(krow["Calls"])
    if a.burst:
        bname, bd = burst_durations(a.stats, a.kernel, a.burst)
        if not bd:
            raise SystemExit("no timing burst in the trace")
        rocprof_avg_us = sum(bd) / len(bd) / 1e3
        n_avg = len(bd)
    live_us = b["roofline"]["avg_launch_us"] if b else None
    alg = b["roofline"]["bytes_per_launch"] if b else None
    # the counter runs keep every slot busy: compare their traffic with the
    # algorithmic bytes of a launch over the whole batch
    ref = bu or b
    alg_full = ref["roofline"].get("bytes_per_launch_full_batch") if ref else None

    lines = [f"# {a.title or a.kernel} — {a.workload}", ""]
    if a.commands and os.path.exists(a.commands):
        lines += ["Commands (repository root, one rocprofv3 run each):", "", "```"]
        lines += open(a.commands).read().strip().splitlines()
        lines += ["```", ""]
    lines += ["## Kernel time (rocprofv3 --kernel-trace --stats)", ""]
    if a.burst:
        lines += ["Several streams ran at once in this trace, so kernels overlap and their durations include the "
                  "sharing; per-kernel times without it are in the one-stream profile.", ""]
    lines += ["| kernel | calls | avg µs | min µs | max µs | % time |", "|---|---|---|---|---|---|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:14]:
        lines.append(f"| `{r['Name']}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    lines += ["", "## Live timing vs rocprof", ""]
    if a.burst:
        lines.append(f"* rocprof average of `{krow['Name']}` over the timing bursts (the launches bench.py times: "
                     f"runs of {a.burst} back-to-back dispatches with the GPU to themselves): "
                     f"{rocprof_avg_us:.2f} µs over {n_avg} launches; over all {krow['Calls']} launches "
                     f"{float(krow['AverageNs']) / 1e3:.2f} µs (other streams' kernels run beside those)")
    else:
        lines.append(f"* rocprof average of `{krow['Name']}`: {rocprof_avg_us:.2f} µs over {n_avg} launches")
    if b:
        lines.append(f"* bench.py HIP-event average of the same kernel in the traced run: {live_us:.2f} µs "
                     f"over {b['roofline']['launches_timed']} sampled launches (back-to-back bursts, so it "
                     f"also holds the tracer's gaps between dispatches)")
        lines.append(f"* algorithmic bytes per launch: {alg} (28 B per p' unknown)")
        lines.append(f"* achieved (bench, events): {b['roofline']['achieved']:.1f} GB/s = "
                     f"{100 * b['roofline']['frac']:.1f} % of 8000 GB/s; "
                     f"with the rocprof average: {alg / (rocprof_avg_us * 1e3):.1f} GB/s")
    if bu:
        r = bu["roofline"]
        lines.append(f"* the same command without the profiler: {r['avg_launch_us']:.2f} µs over "
                     f"{r['launches_timed']} sampled launches, {r['achieved']:.1f} GB/s = {100 * r['frac']:.1f} % "
                     f"of 8000 GB/s (tiling {bu['config'].get('tiling')})")
    if counters:
        lines += ["", "## HBM traffic (rocprofv3 --pmc, one counter per run)", ""]
        lines.append(f"* calibration: {a.copy_mib} MiB device-to-device copy reads FETCH_SIZE {f_copy:.0f} KiB "
                     f"and WRITE_SIZE {w_copy:.0f} KiB -> factors FETCH x{f_factor:.3f}, WRITE x{w_factor:.3f}")
        lines.append(f"* `{kname}`: FETCH_SIZE {f_avg:.0f} KiB/dispatch ({len(kf)} dispatches), "
                     f"WRITE_SIZE {w_avg:.0f} KiB/dispatch ({len(kw)} dispatches)")
        lines.append(f"* calibrated traffic per launch: {traffic / 1e6:.1f} MB"
                     + (f" = {traffic / alg_full:.2f} x the algorithmic bytes of a launch over the full batch "
                        f"({alg_full} B)" if alg_full else ""))
    open(a.out + ".md", "w").write("\n".join(lines) + "\n")
    if counters:
        json.dump({"kernel": kname, "workload": a.workload, "fetch_kib": f_avg, "write_kib": w_avg,
                   "fetch_factor": f_factor, "write_factor": w_factor, "traffic_bytes": traffic,
                   "rocprof_avg_us": rocprof_avg_us, "bench_avg_us": live_us, "bytes_per_launch": alg,
                   "bench_unprofiled_avg_us": bu["roofline"]["avg_launch_us"] if bu else None},
                  open(a.out + ".json", "w"), indent=1)
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    sys.exit(main())
