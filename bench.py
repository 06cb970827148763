#!/usr/bin/env python3
"""Converged solves/s of the SIMPLE backend on one MI355X.

Workload: the 1024 x 512 channel-network case (seeded networks of 6-10
Manhattan/diagonal channels 8-16 cells wide, 2.5D model with Ht = 200 um,
SOU convection, pseudo-CFL 0.5 -> 5 with SER, inlet 0.2 m/s, water),
synthetic geometry generated in-process.

One step = serving a stream of R solve requests (seeds 0..7 in turn, all
arrived at the start) through the library's request loop (cp_serve) with
continuous batching on B slots: mixed precision (FP32 stencil / Jacobi /
multigrid-preconditioned CG with FP64 residuals, refinement and
reductions), one persistent launch per sub-step, the outer iteration
replayed as a hipGraph, finished solves evicted and replaced at once. A
solve counts when the outer stop test max(residU, residV, residMass) <= eps
holds (mass residual over the p' columns) AND its full-system total-pressure
drop is within 1e-4 of the FP64 oracle's converged value
(tests/golden/network_1024x512_dp.json).

Reported beside the value:
  * the plain FP64 single-geometry path (the drop-in: reference algorithm
    and inner budgets, Jacobi-preconditioned CG, polled) on seed 0 with the
    same stop test;
  * p50/p99 latency (admission to completion, and arrival to completion);
  * roofline of the PCG stencil kernel k_pcg_q: bursts of back-to-back
    launches timed with HIP events on the batch stream inside the timed
    region (the kernel recomputes its outputs from unchanged inputs, so a
    burst leaves every solve as it was); the HBM traffic per launch comes
    from the committed rocprofv3 byte-counter profile of this workload;
  * the CPU baseline: the oracle in float arithmetic (the reference's
    arithmetic and algorithm) on the box's host cores for a bounded number
    of outer iterations, scaled to solves/s by the outer-iteration count the
    reference algorithm needs on the GPU.

    python bench.py [--steps K] [--warmup W] [--batch B] [--requests R]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "coldplate-solver_amd"))

import cpbind as cb  # noqa: E402
import workloads as wl  # noqa: E402

METRIC = "converged solves/s + p50/p99 latency at Δp err≤1e-4; PCG HBM GB/s % peak"
PEAK_HBM_GBS = 8000.0
DP_TOL = 1e-4
PCG_Q_BYTES_FP32 = 28  # per p' unknown: read r, aP, p_old, cE, cN; write p_new, q (4 B each)
PROBE_LAUNCHES = 16    # k_pcg_q launches per burst (cp_serve)
GAMMA_MIN = 1.0 - 1e-6  # reference plane row rule (binary geometry)
GEOMS: list = []
TRAFFIC = REPO / "profiles" / "k_pcg_q_network_1024x512_b8.json"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=None,
                    help="slots of the request loop per stream (default: 8; stream workload: 16 over its streams)")
    ap.add_argument("--streams", type=int, default=None,
                    help="batches served at once (cp_serve_pool), each on its own HIP stream and host thread "
                         "(default: 3; stream workload: 2)")
    ap.add_argument("--requests", type=int, default=None,
                    help="solve requests per step (default: 24; stream workload: 120)")
    ap.add_argument("--eps", type=float, default=1e-7)
    ap.add_argument("--jacobi", type=int, default=3, help="Jacobi sweeps per momentum solve")
    ap.add_argument("--pcg", type=int, default=4, help="CG iterations per refinement cycle")
    ap.add_argument("--cycles", type=int, default=1, help="FP64 refinement cycles of p'")
    ap.add_argument("--precond", choices=["amg", "jacobi"], default="amg",
                    help="pressure-CG preconditioner of the served form")
    ap.add_argument("--check", type=int, default=10, help="outer iterations per graph launch between polls")
    ap.add_argument("--workload", choices=["network", "batch64", "stream", "gamma2048"], default="network",
                    help="network: the headline line (1024x512 networks); batch64, stream, gamma2048: the "
                         "batched, served and 2048^2 gamma-field configurations, reported on their own line")
    ap.add_argument("--rate", type=float, default=2.0, help="stream: Poisson arrival rate (requests/s)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--no-probe", action="store_true",
                    help="no k_pcg_q timing bursts")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run 3 device-to-device copies of MIB MiB (byte-counter calibration)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tiling (no autotuning pass)")
    ap.add_argument("--tiling", default="", metavar="R,MG",
                    help="fixed tiling instead of autotuning: rows per wave R and finest multigrid rows MG "
                         "(counter runs use the tiling the autotuner chose)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the step's kernels directly instead of replaying the captured graph")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="stop each batch after this many outer iterations (profiling runs only; "
                         "such solves do not count as converged)")
    return ap.parse_args()


def references(B: int) -> list[float]:
    d = json.loads((REPO / "tests" / "golden" / "network_1024x512_dp.json").read_text())
    out = []
    for s in range(B):
        e = d["seeds"].get(str(s))
        if e is None:
            raise SystemExit(f"no reference pressure drop for seed {s}; run tests/golden/make_network_refs.py")
        out.append(e.get("dp_oracle", e["dp_fp64_gpu"]))
    return out


def n_pressure(M: int, N: int) -> int:
    return (M - 1) * (N - 2)


def solve_stream(bs: list, a, probe: bool):
    """Serve one stream of requests over the batches `bs` (one: cp_serve;
    several: cp_serve_pool, one HIP stream each); returns per-request
    results. Timing bursts run on the first batch, with the GPU to itself."""
    reqs = []
    for k in range(a.requests):
        p, ct = GEOMS[k % len(GEOMS)]
        q = wl.fast_mode(p, a.jacobi, a.pcg, a.cycles)
        q.epsilon = a.eps
        q.pTol = 1e-12
        q.pressurePrecond = cb.CP_PRECOND_AMG if a.precond == "amg" else cb.CP_PRECOND_JACOBI
        if a.profile_steps:
            q.maxIterations = a.profile_steps
        reqs.append((q, ct, 0.0, GAMMA_MIN))
    for k, b in enumerate(bs):
        b.probe(cb.K_PCG_Q if (probe and k == 0) else -1)
    t0 = time.perf_counter()
    res = bs[0].serve(reqs, a.check) if len(bs) == 1 else cb.serve_pool(bs, reqs, a.check)
    elapsed = time.perf_counter() - t0
    ms, unknowns = bs[0].probe_read()
    return dict(elapsed=elapsed, res=res, seeds=[k % len(GEOMS) for k in range(a.requests)], ms=ms,
                unknowns=unknowns)


def dropin_baseline(a, ref0: float):
    """Plain FP64 single-geometry path (reference algorithm and inner budgets,
    polled) on seed 0. It steps until its stop test holds at eps and, when
    its pressure drop is not yet within DP_TOL there, keeps stepping (checking
    every 100 outer iterations) until it is: the time to the first converged
    solve at the required accuracy, a lower bound for this path."""
    p, ct = wl.network_1024x512(0)
    p.massResidualScope = 1
    p.epsilon = 1e-13  # the loop below decides when to stop
    g = cb.Solver(p, ct)
    t0 = time.perf_counter()
    it, err, reached = 0, 1.0, None
    while it < 200000:
        st = g.step()
        it = st["iter"]
        if reached is None and st["error"] <= a.eps:
            reached = it
        if reached is not None and (it == reached or it % 100 == 0):
            err = abs(g.full_dp() / ref0 - 1)
            if err <= DP_TOL:
                break
        if st["converged"]:
            break
    dt = time.perf_counter() - t0
    g.close()
    return dict(seconds=dt, iters=it, iters_at_eps=reached, dp_err=err)


def cpu_baseline(a, outer_iters: int):
    """Oracle in float arithmetic (reference arithmetic), seed 0, bounded."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_py as op  # CPU baseline leg only

    cores = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    p, ct = wl.network_1024x512(0)
    p.massResidualScope = 1
    p.epsilon = a.eps
    o = op.Oracle(p, ct, fp32=True)
    n_done, t_total = 0, 0.0
    while t_total < a.cpu_seconds and n_done < outer_iters:
        n, st, w = o.run(10)
        n_done += n
        t_total += w
        if n < 10:
            break
    per_iter = t_total / max(n_done, 1)
    return dict(value=1.0 / (per_iter * outer_iters), unit="converged solves/s", cores=int(os.environ["OMP_NUM_THREADS"]),
                kind="port", sample=f"float oracle, seed 0, first {n_done} outer iterations in {t_total:.1f} s "
                f"({per_iter * 1e3:.1f} ms/iteration), scaled by the {outer_iters} outer iterations the "
                f"reference algorithm needs on this geometry")


def copy_calibration(mib: int) -> None:
    """Device-to-device copy of a known size (hipMemcpy), so that HBM byte
    counters read under rocprofv3 --pmc can be scaled against it."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy calibration: 3 x {mib} MiB device-to-device")


def served_params(p, a, eps):
    q = wl.fast_mode(p, a.jacobi, a.pcg, a.cycles)
    q.pressurePrecond = cb.CP_PRECOND_AMG if a.precond == "amg" else cb.CP_PRECOND_JACOBI
    q.epsilon = eps
    q.pTol = 1e-12
    q.maxIterations = 20000
    return q


def pcg_roofline(b: cb.Batch) -> dict:
    ms, un = b.probe_read()
    if len(ms) == 0:
        return None
    t = float(np.sum(ms)) * PROBE_LAUNCHES
    byt = float(np.sum(un)) * PCG_Q_BYTES_FP32 * PROBE_LAUNCHES
    ach = byt / t / 1e6
    return {"bound": "hbm", "kernel": "k_pcg_q<float>", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None, "avg_launch_us": round(t / len(ms) / PROBE_LAUNCHES * 1e3, 2),
            "bytes_per_launch": int(byt / len(ms) / PROBE_LAUNCHES), "launches_timed": int(len(ms) * PROBE_LAUNCHES)}


def other_workload(a) -> dict:
    """The batched (64 slots), served (Poisson stream) and 2048^2 gamma-field
    configurations; reported on their own line, never as the headline."""
    base = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32+f64", "data": "synthetic", "steps": 1, "warmup": 0}
    if a.workload == "batch64":
        geoms = wl.batch_64()
        eps = 1e-6
        reqs = [(served_params(p, a, eps), ct, 0.0, 0.5 if k % 2 else GAMMA_MIN) for k, (p, ct) in enumerate(geoms)]
        tmp = cb.Batch(cb.CP_PREC_MIXED, 1 << 20, 1)
        need = sum(tmp.footprint(q.M, q.N) for q, _, _, _ in reqs)
        tmp.close()
        b = cb.Batch(cb.CP_PREC_MIXED, need + (256 << 20), 64)
        b.probe(cb.K_PCG_Q)
        t0 = time.perf_counter()
        res = b.serve(reqs, a.check)
        dt = time.perf_counter() - t0
        conv = sum(r["status"] == cb.SLOT_CONVERGED for r in res)
        lat = [r["tDone"] - r["tAdmit"] for r in res]
        out = dict(base, value=round(conv / dt, 4), unit="converged solves/s", ms_per_step=round(dt * 1e3, 1),
                   config={"workload": "batch64", "geometries": 64, "sizes": "log-uniform 256x128 .. 1024x512",
                           "mix": "32 networks + 32 gamma fields", "stop_test": f"error <= {eps:g}, cap 20000",
                           "slots": 64},
                   converged=conv, capped=len(res) - conv,
                   latency_s={"p50": round(float(np.percentile(lat, 50)), 3),
                              "p99": round(float(np.percentile(lat, 99)), 3)},
                   outer_iterations={"mean": round(float(np.mean([r["iter"] for r in res])), 1),
                                     "max": int(max(r["iter"] for r in res))},
                   roofline=pcg_roofline(b), cpu_baseline=None)
        b.close()
        return out
    if a.workload == "stream":
        n = a.requests or 120
        specs = wl.stream_2000(n, 5)
        arr = wl.poisson_arrivals(n, a.rate)
        eps = 1e-6
        reqs = []
        for r, t in zip(specs, arr):
            p, ct = wl.build_request(r)
            reqs.append((served_params(p, a, eps), ct, float(t), 0.5 if r["kind"] == 2 else GAMMA_MIN))
        nst = a.streams or 2
        per = a.batch or max(1, 16 // nst)
        bs = [cb.Batch(cb.CP_PREC_MIXED, 24 << 30, per) for _ in range(nst)]
        b = bs[0]
        b.probe(cb.K_PCG_Q)
        t0 = time.perf_counter()
        res = b.serve(reqs, a.check) if nst == 1 else cb.serve_pool(bs, reqs, a.check)
        dt = time.perf_counter() - t0
        conv = sum(r["status"] == cb.SLOT_CONVERGED for r in res)
        lat = [r["tDone"] - t for r, t in zip(res, arr)]
        queue = [r["tAdmit"] - t for r, t in zip(res, arr)]
        out = dict(base, value=round(conv / dt, 4), unit="converged solves/s", ms_per_step=round(dt * 1e3, 1),
                   config={"workload": "stream", "requests": n, "arrivals": f"Poisson, {a.rate:g}/s",
                           "sizes": "log-uniform 128^2 .. 2048x1024", "mix": "straight / network / gamma, 2.5D 50%",
                           "streams": nst, "slots_per_stream": per, "stop_test": f"error <= {eps:g}, cap 20000"},
                   converged=conv, capped=len(res) - conv,
                   latency_s={"p50": round(float(np.percentile(lat, 50)), 3),
                              "p99": round(float(np.percentile(lat, 99)), 3), "basis": "arrival to completion",
                              "queue_p50": round(float(np.percentile(queue, 50)), 3)},
                   roofline=pcg_roofline(b), cpu_baseline=None)
        for x in bs:
            x.close()
        return out
    # gamma2048: mixed against the FP64 device path (pinned to the oracle on smaller grids)
    eps = 1e-5
    runs = {}
    for prec in (cb.CP_PREC_MIXED, cb.CP_PREC_FP64):
        p, ct = wl.gamma_2048(prec)
        if prec == cb.CP_PREC_MIXED:
            q = served_params(p, a, eps)
        else:
            q = p.copy()
            q.pressurePrecond = cb.CP_PRECOND_AMG
            q.maxPressureIter = a.pcg
            q.maxMomentumIter = a.jacobi
            q.momentumTol = 0.0
            q.pTol = 1e-12
            q.massResidualScope = 1
            q.epsilon = eps
            q.maxIterations = 20000
        b = cb.Batch(prec, 8 << 30, 1)
        t0 = time.perf_counter()
        r = b.serve([(q, ct, 0.0, 0.5)], a.check)[0]
        runs[prec] = dict(r, seconds=time.perf_counter() - t0)
        b.close()
    m, f = runs[cb.CP_PREC_MIXED], runs[cb.CP_PREC_FP64]
    return dict(base, value=round((1.0 if m["status"] == cb.SLOT_CONVERGED else 0.0) / m["seconds"], 4),
                unit="converged solves/s", ms_per_step=round(m["seconds"] * 1e3, 1),
                config={"workload": "gamma2048", "grid": "2048x2048", "alpha_max": 1e9, "two_five_d": True,
                        "stop_test": f"error <= {eps:g}, cap 20000"},
                mixed={k: m[k] for k in ("status", "iter", "error", "dp", "seconds")},
                fp64_device={k: f[k] for k in ("status", "iter", "error", "dp", "seconds")},
                dp_rel_err_vs_fp64=abs(m["dp"] / f["dp"] - 1), roofline=None, cpu_baseline=None)


def main() -> int:
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X; --gpus {a.gpus} is not supported. "
              f"Independent geometries would shard across GPUs with no communication.", file=sys.stderr)
        return 2
    lib = cb.load()
    if lib.cp_device_count() < 1:
        raise SystemExit("no HIP device")
    if a.workload != "network":
        print(json.dumps(other_workload(a)), flush=True)
        return 0
    if a.calibrate_copy:
        copy_calibration(a.calibrate_copy)
    B = a.batch or 8
    a.streams = a.streams or 3
    a.requests = a.requests or 24
    refs = references(8)
    GEOMS[:] = [wl.network_1024x512(s) for s in range(8)]
    # arena: B of the largest geometry's footprint plus a margin
    probe_b = cb.Batch(cb.CP_PREC_MIXED, 1 << 20, 1)
    need = B * max(probe_b.footprint(p.M, p.N) for p, _ in GEOMS)
    probe_b.close()
    if a.streams < 1:
        raise SystemExit("--streams must be at least 1")
    bs = [cb.Batch(cb.CP_PREC_MIXED, need + (64 << 20), B) for _ in range(a.streams)]
    b = bs[0]
    if a.no_graph:
        for x in bs:
            x.set_graph(False)
    tuned = None
    if a.tiling:
        r_, m_ = (int(x) for x in a.tiling.split(","))
        tuned = {"tile_rows": r_, "mg_rows": m_, "wg_per_cu": 16, "fixed": True}
    elif not a.no_tune:
        # autotune the tiling on the workload's own geometries (one batch
        # alone), then start every measured stream from empty batches
        q0 = [served_params(p, a, a.eps) for p, _ in GEOMS]
        slots = [b.admit(q, ct) for q, (_, ct) in zip(q0[:B], GEOMS[:B])]
        tuned = b.autotune(2)
        for s_ in slots:
            b.evict(s_)
        log(f"autotuned tiling: {tuned}")
    if tuned:
        for x in bs:
            x.set_tiling(tuned["tile_rows"], tuned["wg_per_cu"])
            x.set_mg_tiling(tuned["mg_rows"])
    for w in range(a.warmup):
        r = solve_stream(bs, a, probe=False)
        log(f"warmup {w}: {r['elapsed']:.2f}s iters {[x['iter'] for x in r['res']]}")
    runs = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        r = solve_stream(bs, a, probe=not a.no_probe)
        runs.append(r)
        log(f"step {k}: {r['elapsed']:.2f}s iters {[x['iter'] for x in r['res']]} max dp err "
            f"{max(abs(x['dp'] / refs[sd] - 1) for x, sd in zip(r['res'], r['seeds'])):.2e}")
    total = time.perf_counter() - t0
    good = 0
    errs, lat_service, lat_total, iters = [], [], [], []
    for r in runs:
        for x, sd in zip(r["res"], r["seeds"]):
            e = abs(x["dp"] / refs[sd] - 1)
            errs.append(e)
            lat_service.append(x["tDone"] - x["tAdmit"])
            lat_total.append(x["tDone"])
            iters.append(x["iter"])
            if x["status"] == cb.SLOT_CONVERGED and e <= DP_TOL:
                good += 1
    value = good / total
    # roofline of k_pcg_q from the timing bursts
    ms_all, bytes_all = 0.0, 0.0
    n_launch = 0
    for r in runs:
        for ms, un in zip(r["ms"], r["unknowns"]):
            ms_all += ms * PROBE_LAUNCHES
            bytes_all += PROBE_LAUNCHES * un * PCG_Q_BYTES_FP32
            n_launch += PROBE_LAUNCHES
    avg_ms = ms_all / max(n_launch, 1)
    achieved = bytes_all / max(ms_all, 1e-30) / 1e6  # GB/s
    per_launch = bytes_all / max(n_launch, 1)
    traffic = None
    if TRAFFIC.exists() and B == 8:
        # HBM bytes per k_pcg_q launch over the 8 slots, from the committed
        # rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload
        traffic = int(json.loads(TRAFFIC.read_text())["traffic_bytes"])
    drop = None if a.no_dropin else dropin_baseline(a, refs[0])
    if drop:
        log(f"drop-in FP64 seed 0: {drop}")
    cpu = None
    if not a.no_cpu and drop:
        cpu = cpu_baseline(a, drop["iters"])
        log(f"cpu baseline: {cpu}")
    ms_step = total / max(a.steps, 1) * 1e3
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "converged solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "network_1024x512",
            "requests_per_step": a.requests,
            "streams": a.streams,
            "slots_per_stream": B,
            "grid": "1024x512 cells",
            "stop_test": f"max(residU,residV,residMass[p' cols]) <= {a.eps:g} and dp err <= {DP_TOL:g}",
            "inner_budgets": {"jacobi_sweeps": a.jacobi, "cg_per_cycle": a.pcg, "fp64_refine_cycles": a.cycles},
            "pressure_preconditioner": "aggregation-multigrid V-cycle" if a.precond == "amg" else "Jacobi",
            "serving": ("cp_serve_pool: %d batches on their own streams, one first-come-first-served queue, "
                        "continuous batching, all requests arrived at t=0" % a.streams) if a.streams > 1 else
                       "cp_serve continuous batching, all requests arrived at t=0",
            "tiling": tuned if tuned else "default (no autotuning)",
            "precision": "mixed: fp32 stencil/jacobi/cg/v-cycle, fp64 residuals/refinement/reductions",
        },
        "latency_s": {"p50": round(float(np.percentile(lat_service, 50)), 3),
                      "p99": round(float(np.percentile(lat_service, 99)), 3),
                      "basis": "admission to completion",
                      "from_arrival_p50": round(float(np.percentile(lat_total, 50)), 3),
                      "from_arrival_p99": round(float(np.percentile(lat_total, 99)), 3)},
        "dp_rel_err": {"max": float(max(errs)), "median": float(np.median(errs))},
        "outer_iterations": {"mean": round(float(np.mean(iters)), 1), "max": int(max(iters))},
        "plain_fp64_single": None if not drop else {
            "solves_per_s": round(1.0 / drop["seconds"], 4),
            "seconds_per_solve": round(drop["seconds"], 3),
            "outer_iterations": drop["iters"],
            "outer_iterations_at_eps": drop["iters_at_eps"],
            "dp_rel_err": drop["dp_err"],
            "speedup_of_value": round(value * drop["seconds"], 2),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_pcg_q<float>",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
            "avg_launch_us": round(avg_ms * 1e3, 2),
            "bytes_per_launch": int(per_launch),
            "bytes_per_launch_full_batch": int(PCG_Q_BYTES_FP32 * sum(n_pressure(p.M, p.N) for p, _ in GEOMS[:B])),
            "launches_timed": n_launch,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
