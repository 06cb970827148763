"""GPU tests at the sizes of the five configurations:
  * 1024 x 512 channel network, plain FP64 (the drop-in): ten outer
    iterations against the CPU FP64 oracle, with identical inner iteration
    counts at every step;
  * 2048 x 2048 gamma field: from the same state, one mixed-precision outer
    iteration changes the fields as one FP64 outer iteration does, and the
    boundary conditions hold;
  * all 64 geometries of the batched workload in one batch: the slots run
    to the stop test, packed slots are bit-identical alone, and the
    smallest networks and gamma fields reach the FP64 oracle's pressure
    drop (tests/golden/batch64_small_dp.json);
  * a served Poisson stream whose requests span the served workload's sizes
    (128^2 .. 2048 x 1024) and kinds: every result is bit-identical to the
    same request served alone.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

import cpbind as cb
import oracle_py as op
import workloads as wl

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def served(p, eps=1e-6, cap=20000):
    q = wl.fast_mode(p, 3, 4, 1)
    q.pressurePrecond = cb.CP_PRECOND_AMG
    q.epsilon = eps
    q.pTol = 1e-12
    q.maxIterations = cap
    return q


def test_gamma_2048_mixed_step_matches_fp64_step():
    """From the same FP64 state (after 100 FP64 outer iterations), one
    mixed-precision outer iteration with the same inner budgets (3 Jacobi
    sweeps, 4 multigrid-CG iterations) changes the fields as one FP64 outer
    iteration does, to 1e-3 of the change: the difference is FP32 rounding
    in the inner solves. Boundary conditions hold."""
    p, ct = wl.gamma_2048(cb.CP_PREC_FP64)
    q64 = served(p)
    q64.precision = cb.CP_PREC_FP64
    qmx = served(p)
    qmx.refineCycles = 1  # the same 4 CG iterations as the FP64 step
    say("[configs] gamma 2048^2: 100 FP64 steps")
    g = cb.Solver(q64, ct)
    for _ in range(100):
        g.step()
    before = [g.field(f) for f in (cb.F_U, cb.F_V, cb.F_P)]
    say("[configs] gamma 2048^2: one FP64 and one mixed step from that state")
    g.step()
    after64 = [g.field(f) for f in (cb.F_U, cb.F_V, cb.F_P)]
    g.close()
    m = cb.Solver(qmx, ct)
    m.set_fields(*before)
    st = m.step()
    assert st["health"] == 0
    aftermx = [m.field(f) for f in (cb.F_U, cb.F_V, cb.F_P)]
    m.close()
    # p is compared through its differences between neighbouring fluid
    # cells (gamma > 0.5), the gradients that drive the flow. Its level in a
    # fluid pocket that reaches the outlet only through solid is fixed by
    # couplings ~1e-9 of the fluid's, ill-determined in any precision, and
    # multiplies nothing in the velocity correction.
    fl = np.zeros_like(before[2], dtype=bool)
    fl[1:, 1:] = (1.0 - ct) > 0.5  # p(i, j) is cell (i-1, j-1)
    def grads(f):
        gx = (f[:, 1:] - f[:, :-1])[fl[:, 1:] & fl[:, :-1]]
        gy = (f[1:, :] - f[:-1, :])[fl[1:, :] & fl[:-1, :]]
        return np.concatenate([gx, gy])
    views = [lambda f: f.ravel(), lambda f: f.ravel(), grads]
    for name, b0, a64, amx, view in zip(["u", "v", "grad p"], before, after64, aftermx, views):
        assert np.all(np.isfinite(amx))
        change = np.linalg.norm(view(a64) - view(b0))
        diff = np.linalg.norm(view(amx) - view(a64))
        say(f"  {name}: |step change| {change:.3e}, |mixed - fp64| {diff:.3e}")
        assert diff <= 1e-3 * change, (name, diff, change)
    u, v, _ = aftermx
    assert np.all(u[0, :] == 0) and np.all(u[-1, :] == 0)  # no-slip walls
    assert np.all(v[:, 0] == 0)  # inlet v
    np.testing.assert_array_equal(u[:, -1], u[:, -2])  # zero-gradient outlet


def test_network_1024x512_fp64_matches_oracle_step_by_step():
    """The drop-in (plain FP64, reference algorithm and inner budgets) on the
    1024 x 512 network of seed 0: ten outer iterations, each with the
    oracle's Jacobi and CG iteration counts, and fields within 1e-8 of the
    oracle's after the tenth."""
    p, ct = wl.network_1024x512(0)
    say("[configs] 1024 x 512 network, FP64 drop-in vs oracle, 10 steps")
    g = cb.Solver(p, ct)
    o = op.Oracle(p, ct)
    for step in range(10):
        sg, so = g.step(), o.step()
        assert (sg["itU"], sg["itV"], sg["itP"]) == (so["itU"], so["itV"], so["itP"]), (step, sg, so)
        assert abs(sg["error"] - so["error"]) <= 1e-8 * abs(so["error"]), (step, sg["error"], so["error"])
    for f in (cb.F_U, cb.F_V, cb.F_P):
        a, b = g.field(f), o.field(f)
        assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max(), f
    g.close()


def test_batch64_in_one_batch():
    """The 64 geometries of the batched workload (log-uniform 256x128 ..
    1024x512, half channel networks, half gamma fields) admitted into one
    64-slot batch and run by one launch sequence per outer iteration. Every
    slot reaches its stop test (or the cap); three slots are bit-identical
    when run alone; the two smallest networks and the two smallest gamma
    fields, run to a tight stop test in the same batch, have the FP64
    oracle's converged pressure drop to 1e-4. For the networks that is the
    oracle with the reference's inner budgets. The gamma fields' converged
    flow holds some node at the reference's velocity clamp, so the end
    state of the clamped iteration depends on the inner budgets (the FP64
    oracle with the reference's budgets and with the served ones differ by
    2e-4 there, by 1e-10 on the networks); they are held to the FP64 oracle
    of the served algorithm."""
    refs = json.loads((GOLDEN / "batch64_small_dp.json").read_text())["entries"]
    geoms = wl.batch_64()
    reqs = []
    for k, (p, ct) in enumerate(geoms):
        tight = str(k) in refs
        q = served(p, 1e-9 if tight else 1e-6, 20000)
        reqs.append((q, ct, 0.0, 0.5 if k % 2 else 1.0 - 1e-6))
    tmp = cb.Batch(cb.CP_PREC_MIXED, 1 << 20, 1)
    need = sum(tmp.footprint(q.M, q.N) for q, _, _, _ in reqs)
    tmp.close()
    say(f"[configs] batch_64: 64 geometries in one 64-slot batch, {need / 2**30:.1f} GiB")
    b = cb.Batch(cb.CP_PREC_MIXED, need + (64 << 20), 64)
    res = b.serve(reqs, 20)
    b.close()
    assert all(r["rc"] == cb.CP_OK and r["slot"] >= 0 for r in res)
    assert all(r["tAdmit"] < 0.5 for r in res)  # every geometry admitted at once
    conv = sum(r["status"] == cb.SLOT_CONVERGED for r in res)
    say(f"  converged {conv} of 64; capped {[k for k, r in enumerate(res) if r['status'] != cb.SLOT_CONVERGED]}")
    assert conv >= 62
    for k, e in refs.items():
        r = res[int(k)]
        assert r["status"] == cb.SLOT_CONVERGED, (k, r)
        ref = e["dp_oracle_served"] if e["clamped"] else e["dp_oracle"]
        if not e["clamped"]:
            assert abs(e["dp_oracle_served"] / e["dp_oracle"] - 1) <= 1e-8
        err = abs(r["dp"] / ref - 1)
        say(f"  slot {k} ({e['M']} x {e['N']}, clamped {e['clamped']}): dp {r['dp']:.10g}, oracle {ref:.10g}, "
            f"err {err:.2e}")
        assert err <= 1e-4, (k, err)
    for k in (0, 31, 63):
        b = cb.Batch(cb.CP_PREC_MIXED, 1 << 30, 1)
        alone = b.serve([reqs[k]], 20)[0]
        b.close()
        assert (alone["status"], alone["iter"], alone["dp"]) == (res[k]["status"], res[k]["iter"], res[k]["dp"]), k


def test_served_stream_spanning_every_size_equals_alone():
    """Requests of the served workload's generator chosen to span its sizes
    (the smallest and the largest grid of its 2000 requests, and six
    between) and its kinds and 2.5D settings, with Poisson arrivals into
    3 slots of an arena that holds about three of the largest: admissions
    wait for evictions. Each request runs 150 outer iterations; its result
    (status, iterations, pressure drop) is bit-identical to the same request
    served alone."""
    specs = wl.stream_2000(2000, 5)
    order = sorted(range(len(specs)), key=lambda k: specs[k]["M"] * specs[k]["N"])
    pick = sorted({order[int(round(q * (len(order) - 1)))] for q in np.linspace(0, 1, 8)})
    chosen = [specs[k] for k in pick]
    cells = [r["M"] * r["N"] for r in chosen]
    assert min(cells) <= 160 * 160 and max(cells) >= 1900 * 950, cells
    assert {r["kind"] for r in chosen} == {0, 1, 2}
    arr = wl.poisson_arrivals(len(chosen), 20.0, 9)
    reqs = []
    for r, t in zip(chosen, arr):
        p, ct = wl.build_request(r)
        reqs.append((served(p, 1e-12, 150), ct, float(t), 0.5 if r["kind"] == 2 else 1.0 - 1e-6))
    tmp = cb.Batch(cb.CP_PREC_MIXED, 1 << 20, 1)
    big = max(tmp.footprint(q.M, q.N) for q, _, _, _ in reqs)
    tmp.close()
    say(f"[configs] served stream: {len(reqs)} requests, {min(cells)} .. {max(cells)} cells, 3 slots")
    b = cb.Batch(cb.CP_PREC_MIXED, 3 * big + (64 << 20), 3)
    mix = b.serve(reqs, 25)
    b.close()
    waited = sum(r["tAdmit"] - t > 1e-3 for r, t in zip(mix, arr))
    assert waited >= 1
    for k, (r, req) in enumerate(zip(mix, reqs)):
        assert r["rc"] == cb.CP_OK and r["iter"] == 150 and r["tAdmit"] >= req[2] - 1e-9, (k, r)
        b = cb.Batch(cb.CP_PREC_MIXED, big + (64 << 20), 1)
        alone = b.serve([(req[0], req[1], 0.0, req[3])], 25)[0]
        b.close()
        assert (alone["status"], alone["iter"], alone["dp"]) == (r["status"], r["iter"], r["dp"]), k
