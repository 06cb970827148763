#!/usr/bin/env python3
"""Converged solves/s, served p50/p99 latency and PCG bandwidth of the SIMPLE
backend on one MI355X.

Workload of the headline: the 1024 x 512 channel-network case (seeded
networks of 6-10 Manhattan/diagonal channels 8-16 cells wide, 2.5D model with
Ht = 200 um, SOU convection, pseudo-CFL 0.5 -> 5 with SER, inlet 0.2 m/s,
water), synthetic geometry generated in-process, seeds 0..7.

One step = serving R solve requests (seeds 0..7 in turn, R = twice the slots,
all queued at the start) through the library's request loop with continuous
batching: cp_serve_pool over S batches of B slots, each batch on its own HIP
stream, mixed precision (FP32 stencil / Jacobi / multigrid-preconditioned CG,
FP64 residuals, refinement and reductions), one persistent launch per
sub-step, the outer iteration replayed as a hipGraph. Half of the requests
wait for a slot: every finished solve is read out and evicted and the next
request admitted into its space inside the timed region. A solve counts when
the outer stop test max(residU, residV, residMass) <= eps holds (mass
residual over the p' columns) AND its full-system total-pressure drop is
within 1e-4 of the FP64 oracle's converged value
(tests/golden/network_1024x512_dp.json, every seed).

Reported beside the value (one JSON line):
  * latency_s: a served Poisson stream of the same workload at 0.8 x the
    measured solves/s, more requests than slots, p50/p99 from arrival to
    completion, every solve checked against the oracle's pressure drop;
  * stream_config5: a subsample of the served-stream workload (grid sizes
    log-uniform 128^2 .. 2048 x 1024, straight / network / gamma-field
    geometries, 2.5D on or off) at a fixed Poisson rate: p50/p99 latency,
    converged and capped counts;
  * roofline: the PCG stencil kernel k_pcg_q, timed live inside the timed
    steps: bursts of back-to-back launches between HIP events on the first
    batch's stream, each round launching the kernel over the running slots
    of every batch in turn (the rounds rotate over the whole pool's arrays,
    a working set larger than the 256 MB Infinity Cache); the kernel
    recomputes its outputs from unchanged inputs, so a burst leaves every
    solve as it was; traffic = HBM bytes per launch from the committed
    rocprofv3 byte-counter profile of this workload;
  * roofline_dominant: the same for the finest multigrid level's up-sweep
    k_mg_up, which with k_pcg_q holds the largest shares of GPU time (four
    launches of each per outer iteration); roofline_mom_assemble: the same
    for the momentum assembly, the largest single launch;
  * field_l2_err: relative L2 errors of the served form's final u, v and p
    on seed 0 against the device FP64 fixed point of that geometry (whose
    pressure drop is checked against the CPU FP64 oracle's);
  * plain_fp64_single: the plain FP64 single-geometry path (the drop-in:
    reference algorithm and inner budgets, Jacobi-preconditioned CG,
    polled) on every seed with the same stop test;
  * thermal_cht: the downstream conjugate-heat-transfer solve on the
    seed-0 solution (1024 x 512 x 9 cells), its time per solve and the
    roofline of its finest-level z-line sweep;
  * cpu_baseline: the oracle in float arithmetic (the reference's
    arithmetic and algorithm) on the box's physical host cores for a
    bounded number of outer iterations of seed 0, scaled by the
    outer-iteration count the float oracle needs on that seed to bring its
    pressure drop within 1e-4 of the FP64 oracle's.

    python bench.py [--steps K] [--warmup W] [--batch B] [--streams S]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
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
# algorithmic bytes per work unit of the timed kernels (mixed precision)
PCG_Q_BYTES = 28    # per p' unknown: read z, p_old, aP, cE, cN; write p_new, q (4 B each)
MG_UP_BYTES = 25    # per p' unknown, finest multigrid level: read x2, r, aP, cE, cN (4 B each) and the
#                     coarse z (1 B per unknown: one float per 2 x 2 aggregate); write z
MOM_ASM_BYTES = 72  # per grid node: read u, v, p (8 B each) and the Brinkman faces aU, aV (4 B each); write
#                     dE, dN (4 B each) and per system aP, the face fluxes Fe, Fn and r0 (4 B each)
GAMMA_MIN = 1.0 - 1e-6  # reference plane row rule (binary geometry)
GOLDEN = REPO / "tests" / "golden" / "network_1024x512_dp.json"
TRAFFIC = REPO / "profiles" / "counters_network_1024x512.json"
GEOMS: list = []
N_SEEDS = 8
FIELDS0: dict = {}  # the seed-0 FP64 solution (u, v), the thermal leg's input
# thermal leg: finest-level z-line sweep from a nonzero iterate, per cell:
# read b, x and the nine coefficient / Thomas-factor planes, write x (8 B each)
TH_ZLINE_BYTES = 96
TH_Q0 = 1e6  # W/m^2: the geometry generators' 100 W/cm^2 (GeometryGeneratorV1.py:46)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=None, help="slots per batch (default 20; 8 for --workload stream)")
    ap.add_argument("--streams", type=int, default=None,
                    help="batches served at once (cp_serve_pool), each on its own HIP stream (default 4)")
    ap.add_argument("--requests", type=int, default=None,
                    help="solve requests per step (default: twice the slots)")
    ap.add_argument("--eps", type=float, default=1e-7)
    ap.add_argument("--jacobi", type=int, default=3, help="Jacobi sweeps per momentum solve")
    ap.add_argument("--pcg", type=int, default=4, help="CG iterations per refinement cycle")
    ap.add_argument("--cycles", type=int, default=1, help="FP64 refinement cycles of p'")
    ap.add_argument("--precond", choices=["amg", "jacobi"], default="amg",
                    help="pressure-CG preconditioner of the served form")
    ap.add_argument("--check", type=int, default=10, help="outer iterations per graph launch between polls")
    ap.add_argument("--thermal-solves", type=int, default=3, help="timed thermal CHT solves (0: no thermal leg)")
    ap.add_argument("--workload", choices=["network", "batch64", "stream", "gamma2048", "thermal"], default="network",
                    help="network: the headline line (1024x512 networks); batch64, stream, gamma2048: the "
                         "batched, served and 2048^2 gamma-field configurations, reported on their own line")
    ap.add_argument("--rate", type=float, default=2.0, help="served-stream subsample: Poisson arrival rate (1/s)")
    ap.add_argument("--stream-requests", type=int, default=48,
                    help="requests of the served-stream subsample (0: skip it)")
    ap.add_argument("--latency-requests", type=int, default=None,
                    help="requests of the Poisson latency run of the headline workload (default: twice the slots, "
                         "so that more requests than slots pass through; 0: skip it)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--no-probe", action="store_true", help="no timing bursts")
    ap.add_argument("--no-field-check", action="store_false", dest="field_check",
                    help="skip the L2 errors of the final fields against the FP64 fixed point")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run 3 device-to-device copies of MIB MiB (byte-counter calibration)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tiling (no autotuning pass)")
    ap.add_argument("--tiling", default="", metavar="R,MG[,ORDER[,WG]]",
                    help="fixed tiling instead of autotuning: rows per wave R, finest multigrid rows MG, the "
                         "tile-order mask (cp_batch_set_tile_order, default 0) and workgroups per CU (default 16)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the step's kernels directly instead of replaying the captured graph")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="stop each solve after this many outer iterations (profiling runs only; such solves "
                         "do not count as converged)")
    return ap.parse_args()


def golden() -> dict:
    return json.loads(GOLDEN.read_text())["seeds"]


def references() -> list[float]:
    g = golden()
    out = []
    for s in range(N_SEEDS):
        e = g.get(str(s), {})
        if "dp_oracle" not in e:
            raise SystemExit(f"no FP64-oracle pressure drop for seed {s}; run tests/golden/make_network_refs.py "
                             f"--oracle {s}")
        out.append(e["dp_oracle"])
    return out


def served_params(p, a, eps, cap=20000):
    q = wl.fast_mode(p, a.jacobi, a.pcg, a.cycles)
    q.pressurePrecond = cb.CP_PRECOND_AMG if a.precond == "amg" else cb.CP_PRECOND_JACOBI
    q.epsilon = eps
    q.pTol = 1e-12
    q.maxIterations = a.profile_steps if a.profile_steps else cap
    return q


def field_accuracy(a, tuned, refs: list[float]) -> dict:
    """Relative L2 errors of the served form's final u, v and p on seed 0
    against the FP64 fixed point of the same geometry. The served settings
    are solved alone on a one-slot batch (a slot's result is bit-identical
    alone and in the pool); the reference is the device FP64 path (whose
    sub-steps equal the oracle's) run to a tight fixed point: the oracle's
    strong settings (mass residual over the p' columns, epsilon 1e-9, CG
    tolerance 1e-10, at most 300 CG iterations) with the multigrid
    preconditioner, which changes the path to the fixed point, not the point.
    Its pressure drop is checked against the CPU oracle's."""
    p, ct = GEOMS[0]
    q = served_params(p, a, a.eps)
    pb = cb.Batch(cb.CP_PREC_MIXED, 1 << 20, 1)
    need = pb.footprint(p.M, p.N) + (64 << 20)
    pb.close()
    b = cb.Batch(cb.CP_PREC_MIXED, need, 1)
    if tuned:
        b.set_tiling(tuned["tile_rows"], tuned["wg_per_cu"])
        b.set_mg_tiling(tuned["mg_rows"])
        b.set_tile_order(tuned["tile_order"])
    s = b.admit(q, ct)
    while b.poll()[s]["status"] == cb.SLOT_RUNNING:
        b.run(a.check)
    info = b.poll()[s]
    u, v, pp = b.fields(s)
    dp = b.full_dp(s, GAMMA_MIN)
    b.close()
    r = p.copy()
    r.precision = cb.CP_PREC_FP64
    r.massResidualScope = 1
    r.epsilon = 1e-9
    r.pTol = 1e-10
    r.maxPressureIter = 300
    r.pressurePrecond = cb.CP_PRECOND_AMG
    r.maxIterations = 20000
    t0 = time.perf_counter()
    g = cb.Solver(r, ct)
    st = g.run()
    ref = [g.field(f) for f in (cb.F_U, cb.F_V, cb.F_P)]
    dp64 = g.full_dp(GAMMA_MIN)
    g.close()
    l2 = {n: float(np.linalg.norm(x - y) / np.linalg.norm(y)) for n, x, y in zip(("u", "v", "p"), (u, v, pp), ref)}
    return {"seed": 0, "u": l2["u"], "v": l2["v"], "p": l2["p"], "dp_rel_err": abs(dp / dp64 - 1),
            "served_iterations": info["iter"], "served_status": info["status"],
            "reference": "device FP64 fixed point (epsilon 1e-9, CG tolerance 1e-10, multigrid preconditioner), "
                         f"{st['iter']} outer iterations in {time.perf_counter() - t0:.1f} s; its pressure drop "
                         f"against the CPU FP64 oracle's: {abs(dp64 / refs[0] - 1):.1e}",
            "norm": "||x - x_ref||_2 / ||x_ref||_2 over the field's whole staggered array"}


def network_requests(a, n: int, arrivals=None) -> list:
    reqs = []
    for k in range(n):
        p, ct = GEOMS[k % N_SEEDS]
        reqs.append((served_params(p, a, a.eps), ct, 0.0 if arrivals is None else float(arrivals[k]), GAMMA_MIN))
    return reqs


def serve(bs: list, reqs: list, a, probe: bool):
    """Serve `reqs` over the batches `bs` (cp_serve_pool, one HIP stream
    each); the timing bursts run from the first batch, rotating over all."""
    for k, b in enumerate(bs):
        b.probe((cb.K_PCG_Q, cb.K_MOM_ASM, cb.K_MG_UP) if (probe and k == 0) else ())
    t0 = time.perf_counter()
    res = bs[0].serve(reqs, a.check) if len(bs) == 1 else cb.serve_pool(bs, reqs, a.check)
    elapsed = time.perf_counter() - t0
    return dict(elapsed=elapsed, res=res, bursts=bs[0].probe_read())


def roofline_of(runs: list, kernel: int, unit_bytes: int, name: str, working_set: float, traffic,
                full_units_per_launch: float | None = None) -> dict | None:
    """Algorithmic bytes over event-timed burst time. With
    full_units_per_launch, the figure comes from the bursts taken while every
    slot of the pool was running (each launch covers a full batch, the
    workload's designed shape), and the figure over every burst, including
    those of the step's tail where batches run partly filled, is reported
    beside it."""
    recs = [x for r in runs for x in r["bursts"] if x["kernel"] == kernel]
    if not recs:
        return None

    def rate(rs):
        ms = sum(x["ms"] for x in rs)
        launches = sum(x["launches"] for x in rs)
        units = sum(x["units"] for x in rs)
        return ms, launches, units, units * unit_bytes / (ms * 1e-3) / 1e9

    full = [x for x in recs if full_units_per_launch and
            abs(x["units"] / x["launches"] / full_units_per_launch - 1) < 1e-9]
    ms, launches, units, achieved = rate(full or recs)
    out = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
           "avg_launch_us": round(ms / launches * 1e3, 2), "bytes_per_launch": int(units * unit_bytes / launches),
           "bytes_per_unit": unit_bytes, "launches_timed": int(launches),
           "working_set_bytes_per_round": int(working_set),
           "basis": "bursts with every slot of the pool running" if full else "every burst"}
    if full_units_per_launch:
        ams, al, au, aa = rate(recs)
        out["all_bursts"] = {"achieved": round(aa, 1), "frac": round(aa / PEAK_HBM_GBS, 4),
                             "avg_launch_us": round(ams / al * 1e3, 2), "bytes_per_launch": int(au * unit_bytes / al),
                             "launches_timed": int(al)}
    return out


def dropin_baseline(a, refs: list[float]) -> dict:
    """Plain FP64 single-geometry path (reference algorithm and inner budgets,
    polled) on every seed. Each solve steps until its stop test holds at eps
    and, when its pressure drop is not yet within DP_TOL there, keeps
    stepping (checking every 100 outer iterations) until it is: the time to
    the first converged solve at the required accuracy."""
    per = []
    for s in range(N_SEEDS):
        p, ct = GEOMS[s]
        q = p.copy()
        q.massResidualScope = 1
        q.epsilon = 1e-13  # the loop below decides when to stop
        g = cb.Solver(q, ct)
        t0 = time.perf_counter()
        it, err, reached = 0, 1.0, None
        while it < 100000:
            st = g.step()
            it = st["iter"]
            if reached is None and st["error"] <= a.eps:
                reached = it
            if reached is not None and (it == reached or it % 100 == 0):
                err = abs(g.full_dp() / refs[s] - 1)
                if err <= DP_TOL:
                    break
            if st["converged"] or st["health"]:
                break
        dt = time.perf_counter() - t0
        if s == 0:
            FIELDS0["uv"] = (g.field(cb.F_U), g.field(cb.F_V))
        g.close()
        per.append(dict(seed=s, seconds=round(dt, 3), iters=it, iters_at_eps=reached, dp_err=err))
        log(f"drop-in FP64 seed {s}: {per[-1]}")
    return dict(per_seed=per, seconds=sum(x["seconds"] for x in per))


def physical_cores() -> tuple[int, int]:
    """(physical cores among the CPUs this process may run on, logical CPUs)."""
    cpus = sorted(os.sched_getaffinity(0))
    try:
        out = subprocess.run(["lscpu", "-p=cpu,core,socket"], capture_output=True, text=True, check=True).stdout
        owner = {}
        for line in out.splitlines():
            if line.startswith("#"):
                continue
            c, core, sock = (int(x) for x in line.split(",")[:3])
            owner[c] = (sock, core)
        phys = len({owner[c] for c in cpus if c in owner})
    except (OSError, subprocess.CalledProcessError, ValueError):
        phys = len(cpus)
    return max(phys, 1), len(cpus)


def cpu_baseline(a) -> dict:
    """The oracle in float arithmetic (the reference's arithmetic, algorithm
    and inner budgets), seed 0, for a bounded time on the physical host
    cores; scaled by the outer iterations the float oracle needs on seed 0
    to bring its pressure drop within 1e-4 of the FP64 oracle's."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_py as op  # CPU baseline leg only

    need = golden()["0"].get("iters_float_oracle_dp")
    phys, logical = physical_cores()
    env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(phys, env_threads) if env_threads > 0 else phys
    os.environ["OMP_NUM_THREADS"] = str(threads)
    p, ct = GEOMS[0]
    q = p.copy()
    q.massResidualScope = 1
    q.epsilon = a.eps
    o = op.Oracle(q, ct, fp32=True)
    n_done, t_total = 0, 0.0
    while t_total < a.cpu_seconds:
        n, st, w = o.run(10)
        n_done += n
        t_total += w
        if n < 10:
            break
    per_iter = t_total / max(n_done, 1)
    if not need:
        return dict(value=None, unit="converged solves/s", cores=threads, kind="port",
                    sample=f"float oracle, seed 0, {n_done} outer iterations in {t_total:.1f} s; no committed "
                           f"outer-iteration count of the float oracle to scale by")
    return dict(value=1.0 / (per_iter * need), unit="converged solves/s", cores=threads, kind="port",
                physical_cores_visible=phys, logical_cpus_visible=logical,
                sample=f"float oracle (reference arithmetic, algorithm and inner budgets), seed 0, first {n_done} "
                       f"outer iterations in {t_total:.1f} s ({per_iter * 1e3:.2f} ms/iteration) on {threads} "
                       f"threads, scaled by the {need} outer iterations the float oracle needs on seed 0 until its "
                       f"pressure drop is within 1e-4 of the FP64 oracle's (in single precision its residuals "
                       f"floor above the 1e-7 stop test)")


def thermal_leg(a) -> dict:
    """The downstream conjugate-heat-transfer solve (cp_thermal_solve) on the
    seed-0 solution: 1024 x 512 x 9 cells (the reference's 2 base + 7
    channel layers), 100 W/cm^2 into the base, default rtol 1e-6. Reports
    the wall time per solve (host arrays in and out, device allocation
    included), the device time, and the roofline of its dominant kernel,
    the finest-level z-line sweep, from HIP events around every such sweep.
    CPU comparison: the reference's solver (AMGCL) is not available, so the
    oracle's system is solved directly (SciPy SuperLU) on a 128 x 64 window
    of the same inputs, and the device solves the same window."""
    p, ct = GEOMS[0]
    if "uv" not in FIELDS0:
        q = p.copy()
        q.massResidualScope = 1
        q.epsilon = a.eps
        g = cb.Solver(q, ct)
        g.run()
        FIELDS0["uv"] = (g.field(cb.F_U), g.field(cb.F_V))
        g.close()
    u, v = FIELDS0["uv"]
    gam, ut, vt = cb.thermal_inputs(u, v, ct)

    def prm(Nx, Ny):
        tp = cb.thermal_default_params()
        tp.Nx, tp.Ny, tp.dx, tp.dy, tp.H_t, tp.q0, tp.timing = Nx, Ny, p.hx, p.hy, p.Ht, TH_Q0, 1
        return tp

    tp = prm(p.N, p.M)
    cb.thermal_solve(tp, gam, ut, vt)  # first call: code objects, allocator
    runs = []
    for _ in range(max(a.thermal_solves, 1)):
        t0 = time.perf_counter()
        _, st = cb.thermal_solve(tp, gam, ut, vt)
        runs.append((time.perf_counter() - t0, st))
    n = tp.Nx * tp.Ny * (tp.nz_solid + tp.nz_fluid)
    sweeps = sum(st["sweeps"] for _, st in runs)
    sweep_ms = sum(st["sweep_ms"] * st["sweeps"] for _, st in runs) / max(sweeps, 1)
    achieved = TH_ZLINE_BYTES * n / (sweep_ms * 1e-3) / 1e9
    wall = float(np.median([w for w, _ in runs]))
    st = runs[-1][1]
    out = {"workload": f"thermal CHT on the seed-0 solution, {tp.Nx}x{tp.Ny}x{tp.nz_solid + tp.nz_fluid} cells, "
                       f"q0 {TH_Q0:g} W/m^2, rtol {tp.rtol:g}",
           "solves": len(runs), "ms_per_solve": round(wall * 1e3, 2),
           "device_ms": round(float(np.median([s_["device_ms"] for _, s_ in runs])), 2),
           "iters": st["iters"], "relres": st["relres"], "converged": st["converged"],
           "T_out": round(st["T_out"], 4), "T_base_max": round(st["T_base_max"], 4),
           "imbalance_pct": round(st["imbalance_pct"], 4),
           "roofline": {"bound": "hbm", "kernel": "k_th_zline<9> (finest level, nonzero iterate)",
                        "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                        "avg_launch_us": round(sweep_ms * 1e3, 2), "bytes_per_launch": TH_ZLINE_BYTES * n,
                        "bytes_per_unit": TH_ZLINE_BYTES, "launches_timed": sweeps}}
    if not a.no_cpu:
        sys.path.insert(0, str(REPO / "oracle"))
        import oracle_py as op  # CPU baseline leg only
        import scipy.sparse.linalg as spla

        wn = prm(128, 64)
        win = tuple(np.ascontiguousarray(x[:64, :128]) for x in (gam, ut, vt))
        t0 = time.perf_counter()
        A, b = op.thermal_assemble(wn, *win)
        spla.spsolve(A.tocsc(), b)
        cpu_s = time.perf_counter() - t0
        gpu = []
        for _ in range(3):
            t0 = time.perf_counter()
            cb.thermal_solve(wn, *win)
            gpu.append(time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": round(cpu_s * 1e3, 1), "unit": "ms per solve", "cores": 1, "kind": "port",
                               "sample": "the oracle's system (thermal_oracle.hpp) of a 128x64x9 window of the "
                                         "same inputs, assembled and solved directly by SciPy SuperLU on one "
                                         "host core; AMGCL, the reference's solver, is not in the mount",
                               "device_ms_same_window": round(min(gpu) * 1e3, 2)}
    return out


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


def stream_subsample(a, n: int) -> dict:
    """The served-stream workload's first n requests (seed 5) at a fixed
    Poisson rate, over the same pool shape as the headline; the reference's
    default outer-iteration cap (10000, SIMPLE.h:69)."""
    specs = wl.stream_2000(n, 5)
    arr = wl.poisson_arrivals(n, a.rate)
    reqs = []
    for r, t in zip(specs, arr):
        p, ct = wl.build_request(r)
        reqs.append((served_params(p, a, 1e-6, cap=10000), ct, float(t), 0.5 if r["kind"] == 2 else GAMMA_MIN))
    nst = a.streams
    bs = [cb.Batch(cb.CP_PREC_MIXED, 24 << 30, a.batch) for _ in range(nst)]
    t0 = time.perf_counter()
    res = cb.serve_pool(bs, reqs, a.check) if nst > 1 else bs[0].serve(reqs, a.check)
    dt = time.perf_counter() - t0
    for b in bs:
        b.close()
    conv = [r for r in res if r["status"] == cb.SLOT_CONVERGED]
    lat = [r["tDone"] - t for r, t in zip(res, arr)]
    lat_conv = [r["tDone"] - t for r, t in zip(res, arr) if r["status"] == cb.SLOT_CONVERGED]
    queue = [r["tAdmit"] - t for r, t in zip(res, arr)]
    cells = [s["M"] * s["N"] for s in specs]
    return {"requests": n, "arrivals": f"Poisson, {a.rate:g}/s", "grid_cells": [int(min(cells)), int(max(cells))],
            "sizes": "log-uniform 128^2 .. 2048x1024 (first requests of the seed-5 stream)",
            "mix": "straight / network / gamma field, 2.5D on or off at random",
            "stop_test": "error <= 1e-6, cap 10000 outer iterations", "seconds": round(dt, 2),
            "converged": len(conv), "capped": sum(r["status"] == cb.SLOT_MAXITER for r in res),
            "solves_per_s": round(len(conv) / dt, 4),
            "latency_s": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                          "converged_p50": round(float(np.percentile(lat_conv, 50)), 3) if lat_conv else None,
                          "converged_p99": round(float(np.percentile(lat_conv, 99)), 3) if lat_conv else None,
                          "queue_p50": round(float(np.percentile(queue, 50)), 3),
                          "basis": "arrival to completion"},
            "dp_reference": "none: the FP64 oracle does not finish these grids within the bench; pressure drops are "
                            "checked against the oracle on the 1024x512 workload"}


def other_workload(a) -> dict:
    """The batched (64 slots), served (Poisson stream) and 2048^2 gamma-field
    configurations; reported on their own line, never as the headline."""
    base = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32+f64", "data": "synthetic", "steps": 1, "warmup": 0}
    a.streams = a.streams or 2
    a.batch = a.batch or 8
    if a.workload == "batch64":
        geoms = wl.batch_64()
        eps = 1e-6
        reqs = [(served_params(p, a, eps), ct, 0.0, 0.5 if k % 2 else GAMMA_MIN) for k, (p, ct) in enumerate(geoms)]
        tmp = cb.Batch(cb.CP_PREC_MIXED, 1 << 20, 1)
        need = sum(tmp.footprint(q.M, q.N) for q, _, _, _ in reqs)
        tmp.close()
        b = cb.Batch(cb.CP_PREC_MIXED, need + (256 << 20), 64)
        r = serve([b], reqs, a, probe=not a.no_probe)
        b.close()
        res, dt = r["res"], r["elapsed"]
        conv = sum(x["status"] == cb.SLOT_CONVERGED for x in res)
        lat = [x["tDone"] - x["tAdmit"] for x in res]
        ws = sum(PCG_Q_BYTES * (q.M - 1) * (q.N - 2) for q, _, _, _ in reqs)
        return dict(base, value=round(conv / dt, 4), unit="converged solves/s", ms_per_step=round(dt * 1e3, 1),
                    config={"workload": "batch64", "geometries": 64, "sizes": "log-uniform 256x128 .. 1024x512",
                            "mix": "32 networks + 32 gamma fields", "stop_test": f"error <= {eps:g}, cap 20000",
                            "slots": 64},
                    converged=conv, capped=len(res) - conv,
                    latency_s={"p50": round(float(np.percentile(lat, 50)), 3),
                               "p99": round(float(np.percentile(lat, 99)), 3)},
                    outer_iterations={"mean": round(float(np.mean([x["iter"] for x in res])), 1),
                                      "max": int(max(x["iter"] for x in res))},
                    roofline=roofline_of([r], cb.K_PCG_Q, PCG_Q_BYTES, "k_pcg_q<float>", ws, None),
                    cpu_baseline=None)
    if a.workload == "stream":
        n = a.requests or 120
        out = stream_subsample(a, n)
        return dict(base, value=out["solves_per_s"], unit="converged solves/s", ms_per_step=out["seconds"] * 1e3,
                    config=dict({"workload": "stream", "streams": a.streams, "slots_per_stream": a.batch},
                                **{k: out[k] for k in ("requests", "arrivals", "sizes", "mix", "stop_test")}),
                    converged=out["converged"], capped=out["capped"], latency_s=out["latency_s"], roofline=None,
                    cpu_baseline=None)
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
        raise SystemExit("no gfx950 device")
    if a.workload == "thermal":
        GEOMS[:] = [wl.network_1024x512(0)]
        t = thermal_leg(a)
        print(json.dumps({"metric": "thermal CHT solves/s", "value": round(1e3 / t["ms_per_solve"], 3),
                          "unit": "solves/s", "n_gpus": 1, "steps": t["solves"], "warmup": 1,
                          "ms_per_step": t["ms_per_solve"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": t["workload"]}, "thermal_cht": t,
                          "roofline": t["roofline"], "cpu_baseline": t.get("cpu_baseline")}), flush=True)
        return 0
    if a.workload != "network":
        print(json.dumps(other_workload(a)), flush=True)
        return 0
    if a.calibrate_copy:
        copy_calibration(a.calibrate_copy)
    a.batch = a.batch or 20
    a.streams = a.streams or 4
    if a.streams < 1:
        raise SystemExit("--streams must be at least 1")
    slots = a.batch * a.streams
    a.requests = a.requests or 2 * slots
    if a.latency_requests is None:
        a.latency_requests = 2 * slots
    refs = references()
    GEOMS[:] = [wl.network_1024x512(s) for s in range(N_SEEDS)]
    probe_b = cb.Batch(cb.CP_PREC_MIXED, 1 << 20, 1)
    need = a.batch * max(probe_b.footprint(p.M, p.N) for p, _ in GEOMS)
    probe_b.close()
    bs = [cb.Batch(cb.CP_PREC_MIXED, need + (64 << 20), a.batch) for _ in range(a.streams)]
    if a.no_graph:
        for x in bs:
            x.set_graph(False)
    tuned = None
    if a.tiling:
        f_ = [int(x) for x in a.tiling.split(",")]
        tuned = {"tile_rows": f_[0], "mg_rows": f_[1], "wg_per_cu": f_[3] if len(f_) > 3 else 16,
                 "tile_order": f_[2] if len(f_) > 2 else 0, "fixed": True}
    elif not a.no_tune:
        # autotune the tiling on the workload's own geometries (one batch
        # alone), then start every measured step from empty batches
        b = bs[0]
        q0 = [served_params(p, a, a.eps) for p, _ in GEOMS]
        sl = [b.admit(q, ct) for q, (_, ct) in zip(q0[:a.batch], GEOMS[:a.batch])]
        tuned = b.autotune(2)
        for s_ in sl:
            b.evict(s_)
        log(f"autotuned tiling: {tuned}")
    if tuned:
        for x in bs:
            x.set_tiling(tuned["tile_rows"], tuned["wg_per_cu"])
            x.set_mg_tiling(tuned["mg_rows"])
            x.set_tile_order(tuned["tile_order"])
    for w in range(a.warmup):
        r = serve(bs, network_requests(a, a.requests), a, probe=False)
        log(f"warmup {w}: {r['elapsed']:.2f}s")
    runs = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        r = serve(bs, network_requests(a, a.requests), a, probe=not a.no_probe)
        runs.append(r)
        log(f"step {k}: {r['elapsed']:.2f}s iters {[x['iter'] for x in r['res']]} max dp err "
            f"{max(abs(x['dp'] / refs[i % N_SEEDS] - 1) for i, x in enumerate(r['res'])):.2e}")
    total = time.perf_counter() - t0
    good, errs, iters, service = 0, [], [], []
    for r in runs:
        for i, x in enumerate(r["res"]):
            e = abs(x["dp"] / refs[i % N_SEEDS] - 1)
            errs.append(e)
            iters.append(x["iter"])
            service.append(x["tDone"] - x["tAdmit"])
            good += int(x["status"] == cb.SLOT_CONVERGED and e <= DP_TOL)
    value = good / total
    # served latency at 0.8 x the measured rate, more requests than slots
    latency = None
    if a.latency_requests > 0:
        rate = 0.8 * value
        arr = wl.poisson_arrivals(a.latency_requests, rate, seed=11)
        r = serve(bs, network_requests(a, a.latency_requests, arr), a, probe=False)
        lat = [x["tDone"] - t for x, t in zip(r["res"], arr)]
        queue = [x["tAdmit"] - t for x, t in zip(r["res"], arr)]
        lerr = [abs(x["dp"] / refs[i % N_SEEDS] - 1) for i, x in enumerate(r["res"])]
        lgood = sum(x["status"] == cb.SLOT_CONVERGED and e <= DP_TOL for x, e in zip(r["res"], lerr))
        latency = {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                   "basis": "arrival to completion", "queue_p50": round(float(np.percentile(queue, 50)), 3),
                   "queue_p99": round(float(np.percentile(queue, 99)), 3),
                   "stream": f"{a.latency_requests} requests of this workload, Poisson arrivals at "
                             f"{rate:.2f}/s (0.8 x the measured solves/s) into {slots} slots",
                   "converged_within_dp_tol": int(lgood), "dp_rel_err_max": float(max(lerr)),
                   "seconds": round(r["elapsed"], 2)}
        log(f"latency run: {latency}")
    for x in bs:
        x.close()
    fields = field_accuracy(a, tuned, refs) if a.field_check else None
    if fields:
        log(f"field accuracy: {fields}")
    stream5 = stream_subsample(a, a.stream_requests) if a.stream_requests > 0 else None
    if stream5:
        log(f"served-stream subsample: {stream5}")
    traffic = {}
    t = json.loads(TRAFFIC.read_text()) if TRAFFIC.exists() else {}
    if t and tuned and [tuned["tile_rows"], tuned["mg_rows"]] == t.get("tiling"):
        # HBM bytes per launch, from the committed rocprofv3 FETCH_SIZE /
        # WRITE_SIZE runs of this workload at this batch size and tiling, in
        # the tile order this run used for the kernel's family
        for k, v in t.get("kernels", {}).items():
            bit = (tuned.get("tile_order", 0) >> t["tile_order_bit"][k]) & 1
            e = v["by_launch_order"].get(str(bit))
            if e and t["workload"].startswith(f"network_1024x512, one {a.batch}-slot batch"):
                traffic[k] = int(e["traffic_bytes_per_launch"])
    ws_pcg = PCG_Q_BYTES * slots * (GEOMS[0][0].M - 1) * (GEOMS[0][0].N - 2)
    ws_mom = MOM_ASM_BYTES * slots * (GEOMS[0][0].M + 1) * (GEOMS[0][0].N + 1)
    g0 = GEOMS[0][0]
    roof = roofline_of(runs, cb.K_PCG_Q, PCG_Q_BYTES, "k_pcg_q<float>", ws_pcg, traffic.get("k_pcg_q"),
                       a.batch * (g0.M - 1) * (g0.N - 2))
    roof_dom = roofline_of(runs, cb.K_MG_UP, MG_UP_BYTES, "k_mg_up<float> (finest level)",
                           MG_UP_BYTES * slots * (g0.M - 1) * (g0.N - 2), traffic.get("k_mg_up"),
                           a.batch * (g0.M - 1) * (g0.N - 2))
    roof_mom = roofline_of(runs, cb.K_MOM_ASM, MOM_ASM_BYTES, "k_mom_assemble<float>", ws_mom,
                           traffic.get("k_mom_assemble"), a.batch * (g0.M + 1) * (g0.N + 1))
    drop = None if a.no_dropin else dropin_baseline(a, refs)
    thermal = thermal_leg(a) if a.thermal_solves > 0 else None
    if thermal:
        log(f"thermal leg: {thermal}")
    cpu = None if a.no_cpu else cpu_baseline(a)
    if cpu:
        log(f"cpu baseline: {cpu}")
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "converged solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(total / max(a.steps, 1) * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "network_1024x512",
            "grid": "1024x512 cells",
            "requests_per_step": a.requests,
            "streams": a.streams,
            "slots_per_stream": a.batch,
            "stop_test": f"max(residU,residV,residMass[p' cols]) <= {a.eps:g} and dp err <= {DP_TOL:g} "
                         f"(FP64 oracle, every seed)",
            "inner_budgets": {"jacobi_sweeps": a.jacobi, "cg_per_cycle": a.pcg, "fp64_refine_cycles": a.cycles},
            "pressure_preconditioner": "aggregation-multigrid V-cycle" if a.precond == "amg" else "Jacobi",
            "serving": f"cp_serve_pool: {a.streams} batches of {a.batch} slots on their own streams, one "
                       f"first-come-first-served queue, continuous batching; {a.requests} requests queued at "
                       f"the start of each step, so {max(a.requests - slots, 0)} wait for a slot",
            "tiling": tuned if tuned else "default (no autotuning)",
            "precision": "mixed: fp32 stencil/jacobi/cg/v-cycle, fp64 state, residuals, refinement, reductions",
        },
        "latency_s": latency,
        "service_time_s": {"p50": round(float(np.percentile(service, 50)), 3),
                           "p99": round(float(np.percentile(service, 99)), 3),
                           "basis": "admission to completion in the timed steps"},
        "dp_rel_err": {"max": float(max(errs)), "median": float(np.median(errs)), "reference": "FP64 oracle"},
        "outer_iterations": {"mean": round(float(np.mean(iters)), 1), "max": int(max(iters))},
        "field_l2_err": fields,
        "stream_config5": stream5,
        "plain_fp64_single": None if not drop else {
            "solves_per_s": round(N_SEEDS / drop["seconds"], 4),
            "seconds_all_seeds": round(drop["seconds"], 3),
            "per_seed": drop["per_seed"],
            "speedup_of_value": round(value / (N_SEEDS / drop["seconds"]), 2),
        },
        "roofline": roof,
        "roofline_dominant": roof_dom,
        "roofline_mom_assemble": roof_mom,
        "thermal_cht": thermal,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
