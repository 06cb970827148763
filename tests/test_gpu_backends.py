"""GPU parity of the alternative backends of the reference's seams, through
the C ABI: red-black SOR pressure (pressureSolverType 0,
pressure_solver.cpp:809-909) and explicit momentum (momentumSolverType 0,
iterations.cpp:176-441), alone and combined with the default backends.
FP64 mode: fields within 1e-9 of their scale per outer iteration against the
oracle, identical SOR iteration counts (its stop test compares a maximum,
which no summation order changes). Both backends compute in FP64 whatever
the precision mode, so a mixed-mode slot that uses both equals the FP64 one
bit for bit. In a batch, every slot's result equals the same slot alone.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import cpbind as cb
import oracle_py as op
import workloads as wl

pytestmark = pytest.mark.gpu

FIELDS = [cb.F_U, cb.F_V, cb.F_P, cb.F_USTAR, cb.F_VSTAR, cb.F_PSTAR, cb.F_DE, cb.F_DN]


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def case(kind: str, mom: int, pres: int, precision=cb.CP_PREC_FP64):
    if kind == "straight":
        p, ct = wl.straight_200x50(precision)
    else:
        p, ct = wl.network(48, 96, 5, precision)
    p.momentumSolverType = mom
    p.pressureSolverType = pres
    p.maxPressureIter = 40 if pres == 0 else p.maxPressureIter
    return p, ct


@pytest.mark.parametrize("kind", ["straight", "network"])
@pytest.mark.parametrize("mom,pres", [(0, 0), (0, 1), (1, 0)])
def test_backends_match_oracle_fp64(kind, mom, pres):
    p, ct = case(kind, mom, pres)
    say(f"[backends] {kind} momentum {mom} pressure {pres}")
    g = cb.Solver(p, ct)
    o = op.Oracle(p, ct)
    for step in range(12):
        sg, so = g.step(), o.step()
        for k in ("itU", "itV", "itP", "iter"):
            assert sg[k] == so[k], f"step {step}: {k} {sg[k]} vs {so[k]}"
        for k in ("residU", "residV", "residMass", "error"):
            np.testing.assert_allclose(sg[k], so[k], rtol=1e-8, atol=1e-300, err_msg=f"step {step}: {k}")
        for f in FIELDS:
            e = rel_err(g.field(f), o.field(f))
            assert e <= 1e-9, f"step {step}: field {f} rel err {e:.3e}"
    g.close()


def test_sor_explicit_mixed_equals_fp64():
    p64, ct = case("network", 0, 0, cb.CP_PREC_FP64)
    pmx, _ = case("network", 0, 0, cb.CP_PREC_MIXED)
    a, b = cb.Solver(p64, ct), cb.Solver(pmx, ct)
    for _ in range(20):
        sa, sb = a.step(), b.step()
        assert (sa["itP"], sa["error"]) == (sb["itP"], sb["error"])
    for f in (cb.F_U, cb.F_V, cb.F_P):
        np.testing.assert_array_equal(a.field(f), b.field(f))
    a.close()
    b.close()


def test_mixed_backends_batch_equals_alone():
    """Slots of all four backend pairs in one batch (graph replay), each
    bit-identical to the same slot alone."""
    reqs = []
    for mom, pres in ((0, 0), (1, 1), (0, 1), (1, 0)):
        p, ct = case("network", mom, pres, cb.CP_PREC_MIXED)
        q = wl.fast_mode(p, 3, 4, 1)
        q.momentumSolverType, q.pressureSolverType = mom, pres
        q.maxPressureIter = 20 if pres == 0 else 4
        reqs.append((q, ct))
    say("[backends] batch of four backend pairs")
    b = cb.Batch(cb.CP_PREC_MIXED, 256 << 20, 4)
    slots = [b.admit(q, ct) for q, ct in reqs]
    b.run(60)
    st = b.poll()
    mix = [(st[s]["iter"], st[s]["error"], b.full_dp(s)) for s in slots]
    b.close()
    for k, (q, ct) in enumerate(reqs):
        b = cb.Batch(cb.CP_PREC_MIXED, 128 << 20, 1)
        s = b.admit(q, ct)
        b.run(60)
        st = b.poll()
        alone = (st[s]["iter"], st[s]["error"], b.full_dp(s))
        b.close()
        assert alone == mix[k], (k, alone, mix[k])
        assert np.isfinite(alone[1])


def test_explicit_sor_converges_to_the_implicit_cg_state():
    """Both backend pairs discretise the same equations: run each to a tight
    stop test on the straight channel and compare the pressure drops."""
    out = []
    for mom, pres in ((1, 1), (0, 0)):
        p, ct = wl.straight_200x50()
        p.M, p.N = 20, 60
        ct = wl.geom_straight(20, 60, 5, 14)
        p.momentumSolverType, p.pressureSolverType = mom, pres
        p.massResidualScope = 1
        p.epsilon = 1e-9
        p.maxIterations = 60000
        p.pTol = 1e-12
        p.maxPressureIter = 400
        g = cb.Solver(p, ct)
        st = g.run(500)
        assert st["converged"] == 1, (mom, pres, st)
        out.append(g.full_dp())
        g.close()
    say(f"[backends] converged dp implicit+CG {out[0]:.9g}, explicit+SOR {out[1]:.9g}")
    assert abs(out[1] - out[0]) <= 1e-6 * abs(out[0])
