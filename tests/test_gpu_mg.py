"""GPU tests of the aggregation-multigrid preconditioned pressure solve
(pressurePrecond = CP_PRECOND_AMG) against the oracle's restatement of the
same algorithm, through the C ABI.

Tolerances: FP64 fields within 1e-9 of their scale per step with identical
inner iteration counts (only summation order and contraction differ); the
mixed-precision form to its converged full-system pressure drop within 1e-6
of the FP64 oracle's (the bench's bar is 1e-4) and to its converged u, v
and p within 1e-5 in relative L2 norm.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import cpbind as cb
import oracle_py as op
import workloads as wl

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b) -> float:
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def mg(p, P=None):
    q = p.copy()
    q.pressurePrecond = cb.CP_PRECOND_AMG
    if P is not None:
        q.maxPressureIter = P
    return q


@pytest.mark.parametrize("shape", [(64, 128), (50, 200), (37, 95)], ids=["64x128", "50x200", "37x95"])
def test_mg_fp64_substeps_match_oracle(shape):
    """FP64 drop-in context: every sub-step of 8 outer iterations, with the
    multigrid-preconditioned CG (tight tolerance, so the cap ends it)."""
    M, N = shape
    p, ct = wl.network(M, N, 11)
    p = mg(p, 12)
    p.pTol = 1e-30
    say(f"[mg] fp64 sub-steps {M}x{N}")
    g = cb.Solver(p, ct)
    o = op.Oracle(p, ct)
    for step in range(8):
        sg, so = g.momentum(), o.momentum()
        assert (sg["itU"], sg["itV"]) == (so["itU"], so["itV"])
        sg, so = g.pressure(), o.pressure()
        assert sg["itP"] == so["itP"], f"step {step}: {sg['itP']} vs {so['itP']}"
        e = rel_err(g.field(cb.F_PSTAR), o.field(cb.F_PSTAR))
        assert e <= 1e-9, f"step {step}: p' rel err {e:.2e}"
        g.correct()
        o.correct()
        for f in (cb.F_U, cb.F_V, cb.F_P):
            e = rel_err(g.field(f), o.field(f))
            assert e <= 1e-9, f"step {step}: field {f} rel err {e:.2e}"
    g.close()


def test_mg_fp64_pcg_tolerance_stop():
    """The absolute ||r|| < pTol stop ends the multigrid CG early, with the
    oracle's iteration counts."""
    p, ct = wl.network(48, 96, 4)
    p = mg(p, 50)
    p.pTol = 1e-9
    say("[mg] fp64 tolerance stop")
    g = cb.Solver(p, ct)
    o = op.Oracle(p, ct)
    its = []
    for step in range(10):
        sg, so = g.step(), o.step()
        assert sg["itP"] == so["itP"], f"step {step}: {sg['itP']} vs {so['itP']}"
        its.append(sg["itP"])
    assert max(its) < 50
    for f in (cb.F_U, cb.F_V, cb.F_P):
        assert rel_err(g.field(f), o.field(f)) <= 1e-8
    g.close()


def test_mg_batch_graph_fp64_matches_oracle():
    """Graph-replayed batch with two multigrid slots of different depths and
    one Jacobi slot; every slot matches its own oracle run."""
    cases = [(wl.network(64, 128, 3), True), (wl.network(30, 70, 8), True), (wl.network(40, 90, 2), False)]
    b = cb.Batch(cb.CP_PREC_FP64, 128 << 20, 3)
    slots, refs = [], []
    for (p, ct), use in cases:
        q = mg(p, 10) if use else p.copy()
        q.pTol = 1e-30
        q.maxPressureIter = 10
        slots.append(b.admit(q, ct))
        o = op.Oracle(q, ct)
        o.run(12)
        refs.append(o)
    say("[mg] batch graph fp64")
    b.run(12)
    for s, o in zip(slots, refs):
        u, v, pp = b.fields(s)
        for a, f in ((u, cb.F_U), (v, cb.F_V), (pp, cb.F_P)):
            e = rel_err(a, o.field(f))
            assert e <= 1e-8, f"slot {s} field {f}: {e:.2e}"
    b.close()


def rel_l2(a, b) -> float:
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("kind", ["network", "gamma"])
def test_mg_mixed_converges_to_fp64_oracle(kind):
    """Mixed precision (FP32 V-cycles and CG, FP64 refinement) with four CG
    iterations per cycle converges to the FP64 oracle's fixed point: its
    pressure drop, and its final u, v and p in relative L2 norm over the
    whole staggered arrays."""
    if kind == "network":
        p, ct = wl.network(48, 128, 6, cb.CP_PREC_MIXED)
    else:
        p, ct = wl.gamma_field(48, 96, 3, cb.CP_PREC_MIXED, corr=6.0)
    q = mg(wl.fast_mode(p, 4, 4, 2))
    q.epsilon = 1e-9
    q.pTol = 1e-14
    q.maxIterations = 20000
    say(f"[mg] mixed {kind}")
    b = cb.Batch(cb.CP_PREC_MIXED, 64 << 20, 1)
    s = b.admit(q, ct)
    for _ in range(400):
        b.run(50)
        if b.poll()[s]["status"] != cb.SLOT_RUNNING:
            break
    info = b.poll()[s]
    assert info["status"] == cb.SLOT_CONVERGED, info
    gmin = 0.5 if kind == "gamma" else 1.0 - 1e-6
    dp = b.full_dp(s, gmin)
    r = p.copy()
    r.precision = cb.CP_PREC_FP64
    r = mg(r, 200)
    r.massResidualScope = 1
    r.epsilon = 1e-11
    r.pTol = 1e-14
    r.maxIterations = 20000
    o = op.Oracle(r, ct)
    n, st, _ = o.run(20000)
    assert st["converged"] == 1
    ref = o.full_dp(gmin)
    assert abs(dp / ref - 1) <= 1e-6, (dp, ref, info)
    fields = b.fields(s)
    errs = {n: rel_l2(x, o.field(f)) for n, x, f in zip("uvp", fields, (cb.F_U, cb.F_V, cb.F_P))}
    say(f"[mg] mixed {kind}: relative L2 errors against the oracle {errs}, dp {abs(dp / ref - 1):.2e}")
    for n, e in errs.items():
        assert e <= 1e-5, (n, e, errs)
    b.close()


@pytest.mark.parametrize("precision", [cb.CP_PREC_FP64, cb.CP_PREC_MIXED], ids=["fp64", "mixed"])
def test_mg_alone_equals_mixed_bitwise(precision):
    """A multigrid slot gives bit-identical fields alone and packed with
    slots of other sizes, depths and preconditioners."""
    def prep(p):
        q = wl.fast_mode(p, 4, 4, 2) if precision == cb.CP_PREC_MIXED else p.copy()
        q.precision = precision
        return q
    a = wl.network(64, 128, 21, precision)
    c = wl.gamma_field(40, 96, 5, precision, corr=5.0)
    d = wl.straight_200x50(precision)
    ga = (mg(prep(a[0])), a[1])
    gc = (mg(prep(c[0])), c[1])
    gd = (prep(d[0]), d[1])  # Jacobi
    say(f"[mg] alone vs mixed {precision}")
    alone = cb.Batch(precision, 256 << 20, 4)
    s0 = alone.admit(*ga)
    alone.run(20)
    ref = alone.fields(s0)
    alone.close()
    mixed = cb.Batch(precision, 256 << 20, 4)
    order = [mixed.admit(*gd), mixed.admit(*gc), mixed.admit(*ga)]
    mixed.run(20)
    got = mixed.fields(order[2])
    for x, y in zip(ref, got):
        np.testing.assert_array_equal(x, y)
    mixed.close()


def test_mg_finest_tile_rows_converge_alike():
    """The finest level's tiles of 2, 4 and 8 rows per wave (the autotuner's
    choices in mixed precision) reach the same converged pressure drop: a
    tile shape changes only the grouping of the V-cycle's r.z partial sums."""
    p, ct = wl.network(48, 128, 6, cb.CP_PREC_MIXED)
    q = mg(wl.fast_mode(p, 4, 4, 2))
    q.epsilon = 1e-9
    q.pTol = 1e-14
    q.maxIterations = 20000
    dps = []
    for rows in (2, 4, 8):
        say(f"[mg] finest tile rows {rows}")
        b = cb.Batch(cb.CP_PREC_MIXED, 64 << 20, 1)
        b.set_mg_tiling(rows)
        s = b.admit(q, ct)
        for _ in range(400):
            b.run(50)
            if b.poll()[s]["status"] != cb.SLOT_RUNNING:
                break
        assert b.poll()[s]["status"] == cb.SLOT_CONVERGED
        dps.append(b.full_dp(s))
        b.close()
    for d in dps[1:]:
        assert abs(d / dps[0] - 1) <= 1e-8, dps
