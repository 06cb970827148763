"""Writes tests/golden/batch64_small_dp.json: converged full-system pressure
drops, from the CPU FP64 oracle, of the two smallest channel networks and
the two smallest gamma fields of the 64-geometry batched workload
(workloads.batch_64), for the GPU test that runs the whole batch at once.

Two oracle runs per geometry, both to a fixed point (mass residual over the
p' columns, epsilon 1e-9 and 1e-12, p' by multigrid-preconditioned CG):
  * "dp_oracle": the reference's inner budgets (Jacobi to momentumTol, CG to
    1e-11 within 300 iterations);
  * "dp_oracle_served": the served form's inner budgets (3 Jacobi sweeps, 4
    CG iterations), i.e. the served algorithm in FP64.
Where the reference's velocity clamp |u| <= 3 max(U_in, 0.1)
(momentum_solver.cpp:247, iterations.cpp:104) holds some node of the
converged field at the bound ("clamped"), the clamped iteration's end state
depends on its path and the two differ; elsewhere they agree to round-off.
The plane row rule is the reference's for binary geometries and gamma > 0.5
for gamma fields.

    python tests/golden/make_batch64_refs.py        # CPU, a few minutes
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "coldplate-solver_amd"))
sys.path.insert(0, str(REPO / "oracle"))

OUT = HERE / "batch64_small_dp.json"
PER_KIND = 2


def smallest(geoms):
    order = sorted(range(len(geoms)), key=lambda k: (geoms[k][0].M * geoms[k][0].N, k))
    nets = [k for k in order if k % 2 == 0][:PER_KIND]
    gams = [k for k in order if k % 2 == 1][:PER_KIND]
    return nets + gams


def gamma_min(k: int) -> float:
    return 0.5 if k % 2 else 1.0 - 1e-6  # odd entries are gamma fields


def main() -> int:
    import oracle_py as op
    import workloads as wl

    geoms = wl.batch_64()
    out = {"workload": "batch_64 (seed 4)", "settings": "massResidualScope=1, multigrid-preconditioned CG; "
           "dp_oracle: epsilon 1e-9, pTol 1e-11, maxPressureIter 300; dp_oracle_served: epsilon 1e-12, 3 Jacobi "
           "sweeps, 4 CG iterations", "entries": {}}
    from cpbind import F_U

    for k in smallest(geoms):
        p, ct = geoms[k]
        e = {"M": p.M, "N": p.N, "gamma_min": gamma_min(k)}
        for key, served in (("dp_oracle", False), ("dp_oracle_served", True)):
            q = p.copy()
            q.precision = 0
            q.massResidualScope = 1
            q.pressurePrecond = 1
            q.maxIterations = 60000
            if served:
                q.epsilon = 1e-12
                q.maxMomentumIter, q.momentumTol = 3, 0.0
                q.maxPressureIter, q.pTol = 4, 1e-12
            else:
                q.epsilon = 1e-9
                q.pTol, q.maxPressureIter = 1e-11, 300
            o = op.Oracle(q, ct)
            t = time.time()
            n, st, _ = o.run(60000)
            e[key] = o.full_dp(gamma_min(k))
            e["iters_" + key[3:]] = st["iter"]
            if not served:
                e["clamped"] = bool(abs(o.field(F_U)).max() >= 3 * max(abs(p.U_in), 0.1) * (1 - 1e-12))
            print(k, key, e[key], st["iter"], f"{time.time() - t:.0f}s", flush=True)
        out["entries"][str(k)] = e
        OUT.write_text(json.dumps(out, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
