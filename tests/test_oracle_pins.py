"""Pins of the CPU oracle (no GPU needed).

The reference ships no tests or fixtures, so the oracle is pinned by:
  * golden vectors written by the reference's own compiled solver leaves
    (tests/golden/make_golden.py: jacobi_solver.cpp, pcg_solver.cpp);
  * the reference leaves themselves, live, when oracle/_ref is built;
  * closed forms (plane Poiseuille pressure gradient, Brinkman RAMP alpha),
  * invariants (mirror symmetry, zero inflow, SPD pressure matrix, boundary
    conditions) and dense direct solves on small grids.
"""
from __future__ import annotations

import glob
from pathlib import Path

import numpy as np
import pytest

import oracle_py as op
from cpbind import F_ALPHA, F_DE, F_DN, F_GAMMA, F_P, F_PSTAR, F_U, F_USTAR, F_V, F_VSTAR

GOLDEN = Path(__file__).resolve().parent / "golden"


def straight_params(M=50, N=200, U=0.1):
    p = op.default_params()
    p.M, p.N = M, N
    p.hx = p.hy = 5e-5
    p.rho, p.mu = 1000.0, 1e-3
    p.U_in = U
    p.uvAlpha, p.pAlpha = 0.7, 0.3
    p.epsilon = 1e-6
    p.maxIterations = 100000
    p.usePressureDropConvergence = 0  # residual stop test only
    return p


def straight_geom(M, N, r0, r1):
    ct = np.ones((M, N))
    ct[r0:r1 + 1, :] = 0.0
    return ct


# ---------------------------------------------------------------------------
# golden vectors from the reference's compiled solvers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("path", sorted(glob.glob(str(GOLDEN / "momentum_*.npz"))), ids=lambda p: Path(p).stem)
def test_jacobi_matches_reference_fixture(path):
    d = np.load(path)
    it, x = op.jacobi(d["row_ptr"], d["col_idx"], d["values"], d["b"], d["x0"], d["diag_off"], float(d["tol"]),
                      int(d["max_iter"]), fp32=True)
    assert it == int(d["iters"])
    np.testing.assert_allclose(x, d["x"], rtol=2e-6, atol=1e-7 * max(1e-30, float(np.abs(d["x"]).max())))


@pytest.mark.parametrize("path", sorted(glob.glob(str(GOLDEN / "pressure_*.npz"))), ids=lambda p: Path(p).stem)
def test_pcg_matches_reference_fixture(path):
    d = np.load(path)
    it, x = op.pcg(d["row_ptr"], d["col_idx"], d["values"], d["b"], d["x0"], float(d["tol"]), int(d["max_iter"]),
                   fp32=True)
    assert it == int(d["iters"])
    scale = float(np.abs(d["x"]).max())
    # float PCG over up to 200 iterations: agreement to float round-off growth
    np.testing.assert_allclose(x, d["x"], rtol=0, atol=5e-4 * scale)


@pytest.mark.parametrize("path", [p for p in sorted(glob.glob(str(GOLDEN / "pressure_*.npz"))) if "tight" not in p],
                         ids=lambda p: Path(p).stem)
def test_pcg_fp64_close_to_reference_fixture(path):
    """The FP64 oracle PCG solves the same systems; with the reference's caps
    (tolerances well above float round-off) it takes the same path up to
    float round-off."""
    d = np.load(path)
    it, x = op.pcg(d["row_ptr"], d["col_idx"], d["values"].astype(np.float64), d["b"].astype(np.float64),
                   d["x0"], float(d["tol"]), int(d["max_iter"]), fp32=False)
    scale = float(np.abs(d["x"]).max())
    assert abs(it - int(d["iters"])) <= max(2, int(0.05 * int(d["iters"])))
    np.testing.assert_allclose(x, d["x"], rtol=0, atol=2e-3 * scale)


def test_reference_leaves_live():
    """When oracle/_ref is built, the restated solvers reproduce the
    reference's compiled ones on a freshly assembled system."""
    if op.load_ref() is None:
        pytest.skip("oracle/_ref not built")
    o = op.Oracle(straight_params(20, 60), straight_geom(20, 60, 5, 14), fp32=True)
    for _ in range(5):
        o.step()
    o.momentum()
    s = o.system(0)
    do = op.diag_offsets(s["row_ptr"], s["col_idx"])
    a = op.jacobi(s["row_ptr"], s["col_idx"], s["values"], s["rhs"], s["sol"], do, 1e-9, 50, impl="ref")
    b = op.jacobi(s["row_ptr"], s["col_idx"], s["values"], s["rhs"], s["sol"], do, 1e-9, 50, impl="oracle")
    assert a[0] == b[0]
    np.testing.assert_allclose(a[1], b[1], rtol=1e-6, atol=1e-9)
    o.pressure()
    s = o.system(2)
    x0 = np.zeros(s["n"], np.float32)
    a = op.pcg(s["row_ptr"], s["col_idx"], s["values"], s["rhs"], x0, 1e-7, 60, impl="ref")
    b = op.pcg(s["row_ptr"], s["col_idx"], s["values"], s["rhs"], x0, 1e-7, 60, impl="oracle")
    assert a[0] == b[0]
    np.testing.assert_allclose(a[1], b[1], rtol=0, atol=1e-4 * float(np.abs(a[1]).max()))


# ---------------------------------------------------------------------------
# direct solves, matrix properties
# ---------------------------------------------------------------------------
def dense(s):
    n = s["n"]
    A = np.zeros((n, n))
    for r in range(n):
        for k in range(s["row_ptr"][r], s["row_ptr"][r + 1]):
            A[r, s["col_idx"][k]] = s["values"][k]
    return A


def small_oracle(steps=30):
    p = straight_params(12, 16)
    o = op.Oracle(p, straight_geom(12, 16, 3, 8))
    for _ in range(steps):
        o.step()
    return o


def test_pressure_matrix_spd_and_pcg_reaches_direct_solve():
    o = small_oracle()
    o.momentum()
    o.pressure()
    s = o.system(2)
    A = dense(s)
    np.testing.assert_allclose(A, A.T, rtol=0, atol=1e-18)
    assert np.all(np.linalg.eigvalsh(A) > 0)
    off = np.abs(A).sum(1) - 2 * np.abs(np.diag(A))
    assert np.all(off <= 1e-18)  # weakly diagonally dominant
    x_direct = np.linalg.solve(A, s["rhs"])
    it, x = op.pcg(s["row_ptr"], s["col_idx"], s["values"], s["rhs"], np.zeros(s["n"]), 1e-22, 5000, fp32=False)
    # the reference's absolute |pAp| < 1e-20 guard (pressure_solver.cpp:158)
    # ends the iteration at ~1e-6 relative error on this small-magnitude system
    assert it < 5000
    np.testing.assert_allclose(x, x_direct, rtol=1e-5, atol=1e-6 * np.abs(x_direct).max())


def test_jacobi_fixed_point_is_direct_solve():
    o = small_oracle()
    o.momentum()
    for which in (0, 1):
        s = o.system(which)
        A = dense(s)
        x_direct = np.linalg.solve(A, s["rhs"])
        do = op.diag_offsets(s["row_ptr"], s["col_idx"])
        it, x = op.jacobi(s["row_ptr"], s["col_idx"], s["values"], s["rhs"], np.zeros(s["n"]), do, 1e-30, 20000,
                          fp32=False)
        np.testing.assert_allclose(x, x_direct, rtol=1e-9, atol=1e-12)


def test_momentum_matrix_diagonally_dominant():
    o = small_oracle()
    o.momentum()
    for which in (0, 1):
        A = dense(o.system(which))
        d = np.abs(np.diag(A))
        assert np.all(d >= np.abs(A).sum(1) - d - 1e-15)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def test_brinkman_alpha_ramp_closed_form():
    p = straight_params(10, 12)
    p.alphaMax = 1e9
    ct = np.zeros((10, 12))
    ct[3, :] = 1.0   # gamma 0
    ct[4, :] = 0.5   # gamma 0.5
    o = op.Oracle(p, ct)
    a = o.field(F_ALPHA)
    g = o.field(F_GAMMA)
    q = p.brinkmanQ
    ramp = lambda gm: 1e9 * (1 - gm) / (1 + q * gm + 1e-10)
    np.testing.assert_allclose(a[3], ramp(0.0), rtol=1e-15)
    np.testing.assert_allclose(a[4], ramp(0.5), rtol=1e-15)
    np.testing.assert_allclose(a[5], 0.0, atol=0)
    # rows 0 and M-2 forced solid (SIMPLE.cpp:616-625)
    np.testing.assert_array_equal(a[0], 1e9)
    np.testing.assert_array_equal(a[8], 1e9)
    np.testing.assert_array_equal(g[0], 0.0)
    np.testing.assert_array_equal(g[8], 0.0)
    assert np.all(g[9] == 1.0)


@pytest.fixture(scope="module")
def converged_straight():
    # U = 0.01 m/s: Re = 15 and an entrance length of ~1 mm, so the flow is
    # developed over x = 5..8 mm of the 10 mm channel
    p = straight_params(U=0.01)
    p.massResidualScope = 1
    p.epsilon = 1e-10
    p.pTol, p.maxPressureIter = 1e-12, 500
    o = op.Oracle(p, straight_geom(50, 200, 10, 39))
    n, st, _ = o.run(8000)
    return o, st


def test_poiseuille_pressure_gradient(converged_straight):
    """Developed channel flow against the exact solution of the discrete
    equations: the 30 fluid u-rows (11..40) sit between Brinkman-blocked rows
    10 and 41 (u ~ 0 to 4e-6 relative), so u_k = G h^2 k (31-k) / (2 mu),
    k = 1..30, and the flow rate U_in * 30 h fixes G = 2 mu U_in 30 /
    (h^2 sum k(31-k)). The continuum value for H = 30 h is 11% higher."""
    o, st = converged_straight
    assert st["residU"] < 1e-10 and st["residV"] < 1e-10
    h, mu, U = 5e-5, 1e-3, 0.01
    S = sum(k * (31 - k) for k in range(1, 31))
    G = 2 * mu * U * 30 / (h * h * S)
    p = o.field(F_P)
    i, j0, j1 = 25, 100, 160
    dpdx = (p[i, j1] - p[i, j0]) / ((j1 - j0) * h)
    assert abs(-dpdx / G - 1) < 2e-3
    u = o.field(F_U)[11:41, 130]
    prof = G * h * h * np.array([k * (31 - k) for k in range(1, 31)]) / (2 * mu)
    np.testing.assert_allclose(u, prof, rtol=0, atol=2e-3 * U)


def test_inlet_outlet_mass_balance(converged_straight):
    o, _ = converged_straight
    u = o.field(F_U)
    hy, rho = 5e-5, 1000.0
    m_in = rho * hy * u[1:50, 0].sum()
    m_out = rho * hy * u[1:50, -1].sum()
    assert abs(m_out / m_in - 1) < 1e-6
    np.testing.assert_allclose(m_in, rho * hy * 30 * 0.01, rtol=1e-12)


# ---------------------------------------------------------------------------
# invariants
# ---------------------------------------------------------------------------
def test_zero_inflow_stays_zero():
    p = straight_params(20, 40, U=0.0)
    o = op.Oracle(p, straight_geom(20, 40, 5, 14))
    for _ in range(20):
        st = o.step()
    for f in (F_U, F_V, F_P, F_PSTAR):
        assert np.abs(o.field(f)).max() == 0.0
    assert st["error"] == 0.0


def test_mirror_symmetry():
    """Active cell rows are 0..M-2; a channel symmetric about row (M-2)/2
    gives u even, v odd and p even under the mirror, to FP64 round-off."""
    M, N = 30, 60
    p = straight_params(M, N)
    p.convectionScheme = 1
    o = op.Oracle(p, straight_geom(M, N, 8, 20))  # mirror of row r is 28 - r
    for _ in range(150):
        o.step()
    u, v, pp = o.field(F_U), o.field(F_V), o.field(F_P)
    # u rows 1..M-1 <-> cells 0..M-2: u row i mirrors to M - i
    ui = u[1:M]
    np.testing.assert_allclose(ui, ui[::-1], rtol=0, atol=1e-12 * np.abs(u).max())
    # v rows 1..M-2 (faces) mirror to M-1-i with a sign change
    vi = v[1:M - 1]
    np.testing.assert_allclose(vi, -vi[::-1], rtol=0, atol=1e-12 * np.abs(u).max())
    pi = pp[1:M, 1:N]
    np.testing.assert_allclose(pi, pi[::-1], rtol=0, atol=1e-10 * np.abs(pi).max())
    assert np.abs(vi).max() > 0  # the developing flow has a v component


def test_boundary_conditions_after_step():
    M, N = 20, 40
    p = straight_params(M, N)
    ct = straight_geom(M, N, 5, 14)
    o = op.Oracle(p, ct)
    for _ in range(10):
        o.step()
    u, v, pp, us, vs = o.field(F_U), o.field(F_V), o.field(F_P), o.field(F_USTAR), o.field(F_VSTAR)
    g = o.field(F_GAMMA)
    for a in (u, us):
        assert np.all(a[0] == 0) and np.all(a[M] == 0)
        np.testing.assert_array_equal(a[:, N - 1], a[:, N - 2])
        np.testing.assert_array_equal(a[1:M, 0], np.where(g[0:M - 1, 0] > 0.5, 0.1, 0.0))
    for a in (v, vs):
        assert np.all(a[0] == 0) and np.all(a[M - 1] == 0) and np.all(a[:, 0] == 0)
        np.testing.assert_array_equal(a[:, N], a[:, N - 1])
    np.testing.assert_array_equal(pp[0, 1:N], pp[1, 1:N])
    np.testing.assert_array_equal(pp[M, 1:N], pp[M - 1, 1:N])
    np.testing.assert_array_equal(pp[1:M, 0], pp[1:M, 1])
    assert np.all(pp[1:M, N] == 0)
    assert np.all(o.field(F_PSTAR)[1:M, N - 1] == 0)


def test_d_coefficients_closed_form_first_step():
    """At u = v = 0 (first step, FOU, no pseudo-time) aP0 = 2 De + 2 Dn +
    alpha_face vol, so dE = hy / aP0 in fluid faces."""
    M, N = 12, 20
    p = straight_params(M, N)
    o = op.Oracle(p, np.zeros((M, N)))
    o.momentum()
    dE = o.field(F_DE)
    mu, h = 1e-3, 5e-5
    aP0 = 4 * mu  # De = Dn = mu for square cells
    np.testing.assert_allclose(dE[2:M - 2, 2:N - 2], h / aP0, rtol=1e-14)
    dN = o.field(F_DN)
    np.testing.assert_allclose(dN[2:M - 3, 2:N - 2], h / aP0, rtol=1e-14)


def test_plane_sampling_closed_form():
    """Plane metrics on a hand-set field: uniform u, linear p."""
    M, N = 10, 20
    p = straight_params(M, N)
    ct = straight_geom(M, N, 2, 6)
    o = op.Oracle(p, ct)
    u = np.zeros((M + 1, N)); u[:, :] = 0.2
    v = np.zeros((M, N + 1))
    pp = np.zeros((M + 1, N + 1)); pp[:, :] = 100.0 - np.arange(N + 1)[None, :]
    o.set_fields(u, v, pp)
    x = 5.25 * 5e-5  # between cell centres 4.5 and 5.5 -> columns 4, 5, t = 0.75
    m = o.plane(x)
    rows = 5  # cell rows 2..6 are pure fluid
    assert m["valid"] == 1
    np.testing.assert_allclose(m["flowArea"], rows * 5e-5, rtol=1e-14)
    static = 0.25 * pp[3, 5] + 0.75 * pp[3, 6]
    np.testing.assert_allclose(m["avgStatic"], static, rtol=1e-14)
    np.testing.assert_allclose(m["avgDynamic"], 0.5 * 1000 * 0.04, rtol=1e-14)
    np.testing.assert_allclose(m["avgTotal"], static + 20.0, rtol=1e-14)
    np.testing.assert_allclose(m["massFlux"], rows * 1000 * 0.2 * 5e-5, rtol=1e-14)


def test_fp32_baseline_tracks_fp64_oracle():
    p = straight_params(20, 60)
    ct = straight_geom(20, 60, 5, 14)
    a, b = op.Oracle(p, ct), op.Oracle(p, ct, fp32=True)
    for _ in range(50):
        sa, sb = a.step(), b.step()
    ua, ub = a.field(F_U), b.field(F_U)
    assert np.abs(ua - ub).max() < 2e-3 * np.abs(ua).max()
