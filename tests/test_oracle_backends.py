"""Pins of the oracle's alternative backends, on the CPU:
  * red-black SOR on p' (pressure_solver.cpp:809-909, pressureSolverType 0):
    one red+black sweep pair against a numpy restatement of the sweep, the
    converged result against a dense direct solve of the p' system
    (assembled in numpy from dE, dN, u*, v*), the automatic over-relaxation
    factor against its closed form (through the one-sweep pin), and the stop
    test's iteration count;
  * explicit momentum (iterations.cpp:176-441, momentumSolverType 0): one
    update of u* and v* (and dE, dN) from random fields against a numpy
    restatement of the point update (FOU and pseudo-time; the SOU correction
    is the implicit path's, pinned there);
  * the whole method: explicit momentum + SOR and implicit momentum + CG
    discretise the same equations, so both converge to the same fields.
The reference's own tests hold no fixtures for either backend.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import cpbind as cb
import oracle_py as op
import workloads as wl


def small_channel(M=14, N=24, **kw):
    p = wl.base_params(M, N, 5e-5, 0.05)
    p.enableInletRamp = 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p, wl.geom_straight(M, N, 3, M - 4)


def presys(o: op.Oracle, p) -> tuple[np.ndarray, ...]:
    """The p' coefficients and right-hand side from the oracle's fields."""
    M, N = p.M, p.N
    dE, dN = o.field(cb.F_DE), o.field(cb.F_DN)
    us, vs = o.field(cb.F_USTAR), o.field(cb.F_VSTAR)
    rho, hx, hy = p.rho, p.hx, p.hy
    aE = np.zeros((M + 1, N + 1)); aW = np.zeros_like(aE); aN = np.zeros_like(aE); aS = np.zeros_like(aE)
    b = np.zeros_like(aE)
    for i in range(1, M):
        for j in range(1, N):
            aE[i, j] = rho * dE[i, j] * hy if j < N - 1 else 0.0
            aW[i, j] = rho * dE[i, j - 1] * hy if j > 1 else 0.0
            aN[i, j] = rho * dN[i, j] * hx if i < M - 1 else 0.0
            aS[i, j] = rho * dN[i - 1, j] * hx if i > 1 else 0.0
            b[i, j] = (rho * us[i, j - 1] * hy - rho * us[i, j] * hy) + rho * vs[i - 1, j] * hx - rho * vs[i, j] * hx
    return aE, aW, aN, aS, b


def omega_auto(M, N):
    rs = math.cos(math.pi / max(M, N))
    return max(1.0, min(1.95, 2.0 / (1.0 + math.sqrt(1.0 - rs * rs))))


def sor_sweeps_numpy(aE, aW, aN, aS, b, M, N, omega, sweeps):
    ps = np.zeros((M + 1, N + 1))
    for _ in range(sweeps):
        for color in (0, 1):
            for i in range(1, M):
                for j in range(1, N):
                    if (i + j) % 2 != color:
                        continue
                    aP = aE[i, j] + aW[i, j] + aN[i, j] + aS[i, j]
                    if j == N - 1 or aP < 1e-20:
                        ps[i, j] = 0.0
                        continue
                    pE = ps[i, j + 1] if j < N - 1 else 0.0
                    pW = ps[i, j - 1] if j > 1 else 0.0
                    pN = ps[i + 1, j] if i < M - 1 else 0.0
                    pS = ps[i - 1, j] if i > 1 else 0.0
                    gs = (aE[i, j] * pE + aW[i, j] * pW + aN[i, j] * pN + aS[i, j] * pS + b[i, j]) / aP
                    ps[i, j] = min(10000.0, max(-10000.0, ps[i, j] + omega * (gs - ps[i, j])))
    return ps


def prepared(p, ct, steps=3):
    """An oracle after a few implicit outer iterations, then the momentum
    half of the next one (so dE, dN, u*, v* are non-trivial)."""
    q = p.copy()
    o = op.Oracle(q, ct)
    for _ in range(steps):
        o.step()
    o.momentum()
    return o


@pytest.mark.parametrize("omega", [0.0, 1.4])
def test_sor_two_sweeps_match_numpy(omega):
    p, ct = small_channel(pressureSolverType=0, sorOmega=omega, maxPressureIter=2, pTol=0.0)
    o = prepared(p, ct)
    aE, aW, aN, aS, b = presys(o, p)
    st = o.pressure()
    assert st["itP"] == 2
    w = omega if 0.0 < omega < 2.0 else omega_auto(p.M, p.N)
    ref = sor_sweeps_numpy(aE, aW, aN, aS, b, p.M, p.N, w, 2)
    np.testing.assert_allclose(o.field(cb.F_PSTAR), ref, rtol=1e-12, atol=1e-25)


def test_sor_converges_to_direct_solve():
    p, ct = small_channel(pressureSolverType=0, maxPressureIter=20000, pTol=1e-30)
    o = prepared(p, ct)
    aE, aW, aN, aS, b = presys(o, p)
    o.pressure()
    M, N = p.M, p.N
    idx = {(i, j): k for k, (i, j) in enumerate((i, j) for i in range(1, M) for j in range(1, N - 1))}
    A = np.zeros((len(idx), len(idx)))
    rhs = np.zeros(len(idx))
    for (i, j), k in idx.items():
        A[k, k] = aE[i, j] + aW[i, j] + aN[i, j] + aS[i, j]
        rhs[k] = b[i, j]
        for (ii, jj), c in (((i, j + 1), aE[i, j]), ((i, j - 1), aW[i, j]), ((i + 1, j), aN[i, j]),
                            ((i - 1, j), aS[i, j])):
            if (ii, jj) in idx:
                A[k, idx[(ii, jj)]] = -c
    x = np.linalg.solve(A, rhs)
    ps = o.field(cb.F_PSTAR)
    got = np.array([ps[i, j] for (i, j) in idx])
    np.testing.assert_allclose(got, x, rtol=0, atol=1e-9 * np.abs(x).max())
    assert np.all(ps[1:M, N - 1] == 0.0)


def test_sor_stops_on_max_change():
    tol = 1e-6
    p, ct = small_channel(pressureSolverType=0, maxPressureIter=5000, pTol=tol)
    o = prepared(p, ct)
    aE, aW, aN, aS, b = presys(o, p)
    it = o.pressure()["itP"]
    assert 2 < it < 5000
    w = omega_auto(p.M, p.N)
    # the sweep pair before the last still changed some p' by pTol or more,
    # the last one did not
    a = sor_sweeps_numpy(aE, aW, aN, aS, b, p.M, p.N, w, it - 1)
    c = sor_sweeps_numpy(aE, aW, aN, aS, b, p.M, p.N, w, it)
    prev = sor_sweeps_numpy(aE, aW, aN, aS, b, p.M, p.N, w, it - 2)
    assert np.abs(c - a).max() < tol <= np.abs(a - prev).max()
    np.testing.assert_allclose(o.field(cb.F_PSTAR), c, rtol=1e-10, atol=1e-20)


def explicit_numpy(u, v, pp, alpha, prm, uin, pseudo_dt=None):
    M, N = prm.M, prm.N
    hx, hy, rho, mu, al = prm.hx, prm.hy, prm.rho, prm.mu, prm.uvAlpha
    vol = hx * hy
    maxVel = 3 * max(abs(uin), 0.1)
    De, Dn = mu * hy / hx, mu * hx / hy
    cl = lambda x: np.clip(x, -maxVel, maxVel)

    def cell_alpha(ci, cj):
        return alpha[ci, cj] if 0 <= ci < M and 0 <= cj < N else 0.0

    us, vs = u.copy(), v.copy()
    dE, dN = np.zeros_like(u), np.zeros_like(v)
    ru = rv = 0.0
    for i in range(1, M):
        for j in range(1, N - 1):
            ue = cl(0.5 * (u[i, j] + u[i, j + 1])); uw = cl(0.5 * (u[i, j - 1] + u[i, j]))
            vn = cl(0.5 * (v[i, j] + v[i, j + 1])); vsf = cl(0.5 * (v[i - 1, j] + v[i - 1, j + 1]))
            Fe, Fw, Fn, Fs = rho * hy * ue, rho * hy * uw, rho * hx * vn, rho * hx * vsf
            aE, aW = De + max(0.0, -Fe), De + max(0.0, Fw)
            aN, aS = Dn + max(0.0, -Fn), Dn + max(0.0, Fs)
            tr = rho * vol / pseudo_dt if pseudo_dt else 0.0
            aP0 = max(aE + aW + aN + aS + (Fe - Fw + Fn - Fs) + tr
                      + 0.5 * (cell_alpha(i - 1, j - 1) + cell_alpha(i - 1, j)) * vol, 1e-10)
            dE[i, j] = hy / aP0
            Sp = (pp[i, j] - pp[i, j + 1]) * hy + tr * u[i, j] + (1 - al) / al * aP0 * u[i, j]
            nv = cl((aE * u[i, j + 1] + aW * u[i, j - 1] + aN * u[i + 1, j] + aS * u[i - 1, j] + Sp) / (aP0 / al))
            ru = max(ru, abs(nv - u[i, j]))
            us[i, j] = nv
    for i in range(1, M - 1):
        for j in range(1, N):
            ue = cl(0.5 * (u[i, j] + u[i + 1, j])); uw = cl(0.5 * (u[i, j - 1] + u[i + 1, j - 1]))
            vn = cl(0.5 * (v[i, j] + v[i + 1, j])); vsf = cl(0.5 * (v[i - 1, j] + v[i, j]))
            Fe, Fw, Fn, Fs = rho * hy * ue, rho * hy * uw, rho * hx * vn, rho * hx * vsf
            aE, aW = De + max(0.0, -Fe), De + max(0.0, Fw)
            aN, aS = Dn + max(0.0, -Fn), Dn + max(0.0, Fs)
            tr = rho * vol / pseudo_dt if pseudo_dt else 0.0
            aP0 = max(aE + aW + aN + aS + (Fe - Fw + Fn - Fs) + tr
                      + 0.5 * (cell_alpha(i - 1, j - 1) + cell_alpha(i, j - 1)) * vol, 1e-10)
            dN[i, j] = hx / aP0
            Sp = (pp[i, j] - pp[i + 1, j]) * hx + tr * v[i, j] + (1 - al) / al * aP0 * v[i, j]
            nv = cl((aE * v[i, j + 1] + aW * v[i, j - 1] + aN * v[i + 1, j] + aS * v[i - 1, j] + Sp) / (aP0 / al))
            rv = max(rv, abs(nv - v[i, j]))
            vs[i, j] = nv
    return us, vs, dE, dN, ru, rv


@pytest.mark.parametrize("pseudo", [False, True])
def test_explicit_momentum_matches_numpy(pseudo):
    p, ct = small_channel(momentumSolverType=0, convectionScheme=0)
    if pseudo:
        p.enablePseudoTime = 1
        p.useLocalPseudoTime = 0
    o = op.Oracle(p, ct)
    rng = np.random.default_rng(7)
    M, N = p.M, p.N
    u = rng.uniform(-0.05, 0.1, (M + 1, N))
    v = rng.uniform(-0.02, 0.02, (M, N + 1))
    pp = rng.uniform(0.0, 50.0, (M + 1, N + 1))
    o.set_fields(u, v, pp)
    o.set_scalars(0.05, 1.0)
    st = o.momentum()
    alpha = o.field(cb.F_ALPHA)
    # global pseudo step (SIMPLE.cpp:68-70): timeStepMultiplier * N hx / U_in
    dt = p.timeStepMultiplier * p.N * p.hx / max(p.U_in, 1e-6) if pseudo else None
    us, vs, dE, dN, ru, rv = explicit_numpy(u, v, pp, alpha, p, 0.05, dt)
    inner_u = (slice(1, M), slice(1, N - 1))
    inner_v = (slice(1, M - 1), slice(1, N))
    np.testing.assert_allclose(o.field(cb.F_USTAR)[inner_u], us[inner_u], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(o.field(cb.F_VSTAR)[inner_v], vs[inner_v], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(o.field(cb.F_DE)[inner_u], dE[inner_u], rtol=1e-12)
    np.testing.assert_allclose(o.field(cb.F_DN)[inner_v], dN[inner_v], rtol=1e-12)
    assert st["residU"] == pytest.approx(ru, rel=1e-12)
    assert st["residV"] == pytest.approx(rv, rel=1e-12)
    assert (st["itU"], st["itV"]) == (0, 0)


def test_backends_converge_to_the_same_state():
    """Explicit momentum + SOR and implicit momentum + CG solve the same
    discrete equations: run both to a tight stop test and compare."""
    base = dict(massResidualScope=1, epsilon=1e-9, maxIterations=60000, pTol=1e-12, maxPressureIter=400)
    p1, ct = small_channel(M=12, N=20, **base)
    p0, _ = small_channel(M=12, N=20, pressureSolverType=0, momentumSolverType=0, **base)
    o1, o0 = op.Oracle(p1, ct), op.Oracle(p0, ct)
    n1, s1, _ = o1.run(60000)
    n0, s0, _ = o0.run(60000)
    assert s1["converged"] == 1 and s0["converged"] == 1, (n1, s1["error"], n0, s0["error"])
    for f in (cb.F_U, cb.F_V):
        a, b = o1.field(f), o0.field(f)
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(a), f
    fl = (1.0 - ct) > 0.5
    pa, pb = o1.field(cb.F_P)[1:, 1:][fl], o0.field(cb.F_P)[1:, 1:][fl]
    assert np.linalg.norm(pa - pb) <= 1e-5 * np.linalg.norm(pa)
