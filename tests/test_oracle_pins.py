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


# ---------------------------------------------------------------------------
# second-order upwind deferred correction (convection.cpp:53-186)
# ---------------------------------------------------------------------------
def sou_source(p, ct, u, v, which):
    """S_dc of every row of system `which` (0 = U, 1 = V): the difference of
    the right-hand sides assembled with SOU and with FOU on the same state
    (nothing else in a row depends on the scheme)."""
    rhs = {}
    for conv in (0, 1):
        q = p.copy()
        q.convectionScheme = conv
        o = op.Oracle(q, ct)
        o.set_fields(u, v, np.zeros((p.M + 1, p.N + 1)))
        o.momentum()
        s = o.system(which)
        rhs[conv] = s["rhs"]
    # the difference carries the round-off of the full right-hand sides
    return rhs[1] - rhs[0], s["ri"], s["rj"], np.abs(rhs[0]).max()


# Index guards of the two corrections (convection.cpp:53-118 for U,
# 121-186 for V): per face, the condition on (i, j, M, N) under which the
# upwind-upwind node is used, for F >= 0 and for F < 0.
SOU_GUARDS = {
    "U": {"e": (lambda i, j, M, N: j >= 2, lambda i, j, M, N: j + 2 < N),
          "w": (lambda i, j, M, N: j >= 2, lambda i, j, M, N: j + 1 < N),
          "n": (lambda i, j, M, N: i >= 1, lambda i, j, M, N: i + 2 <= M),
          "s": (lambda i, j, M, N: i >= 2, lambda i, j, M, N: i + 1 <= M)},
    "V": {"e": (lambda i, j, M, N: j >= 2, lambda i, j, M, N: j + 2 <= N),
          "w": (lambda i, j, M, N: j >= 2, lambda i, j, M, N: j + 1 <= N),
          "n": (lambda i, j, M, N: i >= 1, lambda i, j, M, N: i + 2 < M),
          "s": (lambda i, j, M, N: i >= 2, lambda i, j, M, N: i + 1 < M)},
}


def sou_linear_closed_form(kind, ri, rj, M, N, rho, hx, hy, uc, vc, phi):
    """S_dc for linear fields. u = uc[0] + uc[1] i + uc[2] j and v likewise:
    on a linear field the SOU face value 1.5 phi_U - 0.5 phi_UU extrapolates
    exactly, so phi_SOU - phi_FOU = +-0.5 (slope along the face normal), + for
    F >= 0 and - for F < 0. The fluxes are F_e,w = rho hy u_face and
    F_n,s = rho hx v_face with face velocities the means of the two nodes
    that straddle the face. S_dc = -F_e d_e + F_w d_w - F_n d_n + F_s d_s."""
    U = lambda i, j: uc[0] + uc[1] * i + uc[2] * j
    V = lambda i, j: vc[0] + vc[1] * i + vc[2] * j
    g = SOU_GUARDS[kind]
    bi, bj = phi[1], phi[2]
    out = np.zeros(len(ri))
    for r, (i, j) in enumerate(zip(ri, rj)):
        if kind == "U":
            ue, uw = (U(i, j) + U(i, j + 1)) / 2, (U(i, j - 1) + U(i, j)) / 2
            vn, vs = (V(i, j) + V(i, j + 1)) / 2, (V(i - 1, j) + V(i - 1, j + 1)) / 2
        else:
            ue, uw = (U(i, j) + U(i + 1, j)) / 2, (U(i, j - 1) + U(i + 1, j - 1)) / 2
            vn, vs = (V(i, j) + V(i + 1, j)) / 2, (V(i - 1, j) + V(i, j)) / 2
        F = {"e": rho * hy * ue, "w": rho * hy * uw, "n": rho * hx * vn, "s": rho * hx * vs}
        slope = {"e": bj, "w": bj, "n": bi, "s": bi}
        sgn = {"e": -1.0, "w": 1.0, "n": -1.0, "s": 1.0}
        S = 0.0
        for f in "ewns":
            pos = F[f] >= 0
            if g[f][0 if pos else 1](i, j, M, N):
                S += sgn[f] * F[f] * (0.5 * slope[f] if pos else -0.5 * slope[f])
        out[r] = S
    return out


@pytest.mark.parametrize("case", [
    # (u coefficients a + b i + c j, v coefficients): all fluxes >= 0, all < 0,
    # and flows that change direction inside the grid
    ((0.05, 0.0, 0.004), (0.02, 0.003, 0.0)),
    ((-0.05, 0.0, -0.004), (-0.02, -0.003, 0.0)),
    ((0.04, 0.006, 0.003), (-0.03, 0.004, 0.002)),
    ((-0.01, -0.004, 0.002), (0.05, -0.006, -0.003)),
], ids=["positive", "negative", "mixed-a", "mixed-b"])
def test_sou_exact_on_linear_fields_with_every_guard(case):
    """On linear u and v the SOU correction of every U and V row has the
    closed form of sou_linear_closed_form, including the rows next to the
    grid edges, where the reference's index guards (asymmetric between U and
    V) drop a face. Swapped 1.5/0.5 weights, a sign slip or a guard moved by
    one row or column fail it."""
    M, N = 14, 18
    p = straight_params(M, N, U=0.1)  # face speeds stay below 3 max(U_in, 0.1)
    p.alphaMax = 1.0  # weak forced wall rows: the rows' sources stay of the size of S_dc
    ct = np.zeros((M, N))
    uc, vc = case
    ii, jj = np.mgrid[0:M + 1, 0:N]
    u = uc[0] + uc[1] * ii + uc[2] * jj
    ii, jj = np.mgrid[0:M, 0:N + 1]
    v = vc[0] + vc[1] * ii + vc[2] * jj
    for which, kind, phi in ((0, "U", uc), (1, "V", vc)):
        S, ri, rj, rhs_scale = sou_source(p, ct, u, v, which)
        ref = sou_linear_closed_form(kind, ri, rj, M, N, p.rho, p.hx, p.hy, uc, vc, phi)
        assert np.abs(ref).max() > 1e-4 * rhs_scale
        np.testing.assert_allclose(S, ref, rtol=0, atol=1e-14 * rhs_scale)
        # every guard row is exercised: the edge rows differ from the interior
        assert np.any(S != S[len(S) // 2])


def test_sou_vanishes_on_constant_fields():
    M, N = 12, 16
    p = straight_params(M, N)
    S, _, _, _ = sou_source(p, np.zeros((M, N)), np.full((M + 1, N), 0.07), np.full((M, N + 1), -0.03), 0)
    assert np.abs(S).max() == 0.0


# ---------------------------------------------------------------------------
# 2.5D depth-averaged sink (momentum_solver.cpp:187-190, SIMPLE.h:205-208)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def converged_sink_channel():
    # 20 fluid cell rows (u rows 6..25) between solid rows, h = 50 um, Ht =
    # 500 um: k = 4.8 * 5 mu / (2 Ht^2) = 4.8e4, 1 / sqrt(k / mu) = 144 um.
    # At U = 0.01 m/s (Re = 10) u and v are x-independent to 1e-11 of U from
    # column 90 on.
    M, N = 32, 120
    p = straight_params(M, N, U=0.01)
    p.enableTwoPointFiveD = 1
    p.Ht = 5e-4
    p.massResidualScope = 1
    p.epsilon = 1e-11
    p.pTol, p.maxPressureIter = 1e-13, 200
    p.pressurePrecond = 1  # multigrid-preconditioned CG: the p' solves converge in a few iterations
    o = op.Oracle(p, straight_geom(M, N, 5, 24))
    n, st, _ = o.run(20000)
    return o, p, st


def test_sink_developed_profile_solves_discrete_equation(converged_sink_channel):
    """In developed flow (v = 0, u independent of x) the U equation with the
    2.5D sink reduces, per row, to mu (u[i+1] - 2 u[i] + u[i-1]) / hy^2 -
    (k + alpha_i) u[i] = -(p[i,j] - p[i,j+1]) / hx with k = 4.8 * 5 mu / (2
    Ht^2) and alpha_i the Brinkman term of the row. The oracle's converged
    centre column equals the direct tridiagonal solve of that equation with
    the oracle's own pressure differences."""
    o, p, st = converged_sink_channel
    assert st["error"] <= 1e-11
    M, N, mu, h = p.M, p.N, p.mu, p.hx
    k = 4.8 * 5 * mu / (2 * p.Ht ** 2)
    j = 90
    u = o.field(F_U)
    pp = o.field(F_P)
    alpha = o.field(F_ALPHA)
    n = M - 1  # u rows 1 .. M-1, u[0] = u[M] = 0
    A = np.zeros((n, n))
    rhs = np.zeros(n)
    for r in range(n):
        i = r + 1
        a_face = 0.5 * (alpha[i - 1, j - 1] + alpha[i - 1, j])
        A[r, r] = 2 * mu / h ** 2 + k + a_face
        if r > 0:
            A[r, r - 1] = -mu / h ** 2
        if r < n - 1:
            A[r, r + 1] = -mu / h ** 2
        rhs[r] = (pp[i, j] - pp[i, j + 1]) / h
    prof = np.linalg.solve(A, rhs)
    np.testing.assert_allclose(u[1:M, j], prof, rtol=0, atol=1e-7 * p.U_in)
    # the sink's strength is what sets the profile: without it (k = 0) the
    # same pressure differences give a different one
    A0 = A - k * np.eye(n)
    assert np.abs(np.linalg.solve(A0, rhs) - u[1:M, j]).max() > 0.1 * p.U_in


def test_sink_pressure_gradient_near_continuum(converged_sink_channel):
    """Continuum bound: between walls H apart, mu u'' - k u = G with mean
    velocity U gives G = -k U / (1 - tanh(lH/2) / (lH/2)), l = sqrt(k / mu).
    The walls of the discrete channel sit at the first blocked u rows (5 and
    26), so H = 21 h; the discrete value is within a few percent."""
    o, p, _ = converged_sink_channel
    mu, h = p.mu, p.hx
    k = 4.8 * 5 * mu / (2 * p.Ht ** 2)
    lam = np.sqrt(k / mu)
    H = 21 * h
    U = p.U_in * 20 / 21  # the inflow of 20 rows spread over H
    G = -k * U / (1 - np.tanh(lam * H / 2) / (lam * H / 2))
    pp = o.field(F_P)
    dpdx = (pp[15, 90] - pp[15, 60]) / (30 * h)
    assert abs(dpdx / G - 1) < 0.05, (dpdx, G)


# ---------------------------------------------------------------------------
# 6/7 convection scaling (momentum_solver.cpp:220) and the pseudo-time term
# (momentum_solver.cpp:231-242): invariants of the converged fixed point
# ---------------------------------------------------------------------------
def obstacle_geom(M, N):
    ct = straight_geom(M, N, 3, M - 5)
    ct[M // 2 - 2:M // 2 + 1, N // 3:N // 3 + 4] = 1.0  # a block in the channel: 2D flow
    return ct


def run_fixed_point(p, ct, steps=30000):
    o = op.Oracle(p, ct)
    first = o.step()
    n, st, _ = o.run(steps)
    assert st["error"] <= p.epsilon, st
    return o, first


def tight(p):
    # fixed points to 1e-11, with p' solved far below the outer residual by
    # multigrid-preconditioned CG (the preconditioner changes no fixed point)
    p.massResidualScope = 1
    p.epsilon = 1e-11
    p.pTol, p.maxPressureIter = 1e-16, 200
    p.pressurePrecond = 1
    return p


def test_convection_scaling_is_a_density_change_of_the_fluxes():
    """The 6/7 factor multiplies the convective fluxes only: with it on at
    density rho, every momentum coefficient equals that of the unscaled
    equations at 6 rho / 7, and the pressure-correction system is the same up
    to its factor rho, which cancels from p'. So both runs reach the same u,
    v, p, and the first iteration's mass imbalance scales by exactly 6/7. A
    factor applied to the diffusion too, or 7/6, fails."""
    M, N = 22, 48
    ct = obstacle_geom(M, N)
    a = tight(straight_params(M, N, U=0.05))
    a.enableConvectionScaling = 1
    b = tight(straight_params(M, N, U=0.05))
    b.rho = a.rho * 6 / 7
    oa, fa = run_fixed_point(a, ct)
    ob, fb = run_fixed_point(b, ct)
    assert abs(fb["residMass"] / fa["residMass"] - 6 / 7) < 1e-12
    for f in (F_U, F_V, F_P):
        x, y = oa.field(f), ob.field(f)
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-8 * np.abs(x).max())
    # the fixed point does depend on the factor: unscaled at rho differs
    c = tight(straight_params(M, N, U=0.05))
    oc, _ = run_fixed_point(c, ct)
    assert np.abs(oc.field(F_U) - oa.field(F_U)).max() > 1e-6 * np.abs(oa.field(F_U)).max()


@pytest.mark.parametrize("local", [1, 0], ids=["local-dt", "global-dt"])
def test_pseudo_time_leaves_the_fixed_point_unchanged(local):
    """The pseudo-time term adds rho vol / dt to the diagonal and rho vol /
    dt times the old velocity to the source; at a fixed point u* = u = uOld,
    so they cancel and the converged fields equal those without pseudo-time
    (SER and line search on, as the reference's defaults). A wrong sign on
    either part moves the fixed point."""
    M, N = 22, 48
    ct = obstacle_geom(M, N)
    a = tight(straight_params(M, N, U=0.05))
    a.convectionScheme = 1
    b = a.copy()
    b.enablePseudoTime = 1
    b.useLocalPseudoTime = local
    oa, _ = run_fixed_point(a, ct)
    ob, _ = run_fixed_point(b, ct)
    for f in (F_U, F_V, F_P):
        x, y = oa.field(f), ob.field(f)
        np.testing.assert_allclose(y, x, rtol=0, atol=1e-8 * np.abs(x).max())


# ---------------------------------------------------------------------------
# pseudo-CFL controller: SER, line search, residual ramp
# (time_control.cpp:43-165 in the order of SIMPLE.cpp:321-327)
# ---------------------------------------------------------------------------
def ser_params():
    p = op.default_params()  # SER on: CFL 0.1 .. 5, smoothing 0.5, ratio 0.7 .. 1.1, from iteration 10
    p.enablePseudoTime = 1
    p.enableCflRamp = 1
    p.pseudoCFLInitial = 0.5
    return p


def test_ser_constant_residual_holds_cfl():
    p = ser_params()
    cfl, ls = op.cfl_sequence(p, np.full(40, 3e-4))
    np.testing.assert_array_equal(cfl, 0.5)
    np.testing.assert_array_equal(ls, 1.0)


def test_ser_waits_for_min_iteration_then_grows_5_percent_per_step_to_the_cap():
    """A residual falling faster than 10 % per step saturates the ratio at
    serMaxIncrease = 1.1; with smoothing 0.5 the CFL grows by 1.05 per step
    until 1.1 CFL passes serCFLMax, then relaxes onto serCFLMax."""
    p = ser_params()
    res = 1e-2 * 0.8 ** np.arange(130)  # stays above the 1e-15 floor of the monitored residual
    cfl, ls = op.cfl_sequence(p, res)
    assert np.all(cfl[:11] == 0.5)  # iterations 0..9 wait; 10 only records the residual
    k = 11
    while 1.1 * cfl[k - 1] <= p.serCFLMax:
        assert cfl[k] == pytest.approx(1.05 * cfl[k - 1], rel=1e-14)
        k += 1
    assert k > 40  # 0.5 * 1.05^n reaches 5 / 1.1 after ~31 steps
    assert np.all(cfl <= p.serCFLMax)
    assert cfl[-1] == pytest.approx(p.serCFLMax, rel=1e-12)
    np.testing.assert_array_equal(ls, 1.0)


def test_ser_rising_residual_cuts_15_percent_per_step_to_the_floor():
    """A residual rising faster than 1 / 0.7 per step saturates the ratio at
    serMaxDecrease = 0.7: the CFL falls by 0.85 per step until 0.7 CFL passes
    serCFLMin, then relaxes onto it."""
    p = ser_params()
    p.pseudoCFLInitial = 4.0
    cfl, ls = op.cfl_sequence(p, 1e-6 * 2.0 ** np.arange(120))
    k = 11
    while 0.7 * cfl[k - 1] >= p.serCFLMin:
        assert cfl[k] == pytest.approx(0.85 * cfl[k - 1], rel=1e-14)
        k += 1
    assert cfl.min() >= p.serCFLMin
    assert cfl[-1] == pytest.approx(p.serCFLMin, rel=1e-12)
    # in the reference's call order SER stores the new residual before the
    # line search compares with it, so with its 1 % tolerance the line search
    # never cuts (a 0.3 cut would show here)
    np.testing.assert_array_equal(ls, 1.0)


def test_line_search_cuts_when_its_tolerance_is_below_one():
    """With lsResidIncreaseTol < 1 the ratio of 1 left by SER counts as a
    rise: each iteration from SER's start cuts the step size by 0.3 (floor
    lsAlphaMin) and the CFL by 0.3 (floor serCFLMin), for lsMaxTries
    consecutive iterations; after that only SER acts."""
    p = ser_params()
    p.pseudoCFLInitial = 4.0
    p.lsResidIncreaseTol = 0.5
    cfl, ls = op.cfl_sequence(p, np.full(20, 1e-3))
    assert np.all(cfl[:10] == 4.0) and np.all(ls[:10] == 1.0)
    # iteration 10 records the residual (ratio 1 afterwards, SER keeps the CFL)
    assert cfl[10] == pytest.approx(4.0 * 0.3) and ls[10] == pytest.approx(0.3)
    assert cfl[11] == pytest.approx(4.0 * 0.09) and ls[11] == pytest.approx(0.1)
    assert cfl[12] == pytest.approx(4.0 * 0.027) and ls[12] == pytest.approx(0.1)
    assert cfl[13] == pytest.approx(p.serCFLMin)  # 4 * 0.0081 floored
    np.testing.assert_array_equal(cfl[14:], cfl[13])  # tries used up, constant residual


def test_residual_ramp_without_ser():
    """SER off: above cflRampStartRes the CFL is the initial one; below it
    the target is min(initial (start / r)^exponent, max), approached with
    weight cflRampSmooth per iteration."""
    p = ser_params()
    p.enableSER = 0
    p.pseudoCFLInitial, p.pseudoCFLMax = 0.2, 30.0
    start = p.cflRampStartRes
    res = np.array([10 * start, 2 * start, start / 10, start / 10, start / 1e9, start / 1e9])
    cfl, _ = op.cfl_sequence(p, res)
    assert cfl[0] == 0.2 and cfl[1] == 0.2
    tgt = min(0.2 * 10 ** p.cflRampExponent, 30.0)
    s = p.cflRampSmooth
    assert cfl[2] == pytest.approx((1 - s) * 0.2 + s * tgt, rel=1e-14)
    assert cfl[3] == pytest.approx((1 - s) * cfl[2] + s * tgt, rel=1e-14)
    assert cfl[4] == pytest.approx((1 - s) * cfl[3] + s * 30.0, rel=1e-14)  # capped target
    p.enablePseudoTime = 0
    np.testing.assert_array_equal(op.cfl_sequence(p, res)[0], 0.2)


# ---------------------------------------------------------------------------
# pressure-drop window exit (SIMPLE.cpp:356-409): special cases
# ---------------------------------------------------------------------------
def dp_params(window=20, ramp=5, pct=1.0, gate=0, slope_window=20, slope_max=0.05):
    p = op.default_params()
    p.usePressureDropConvergence = 1
    p.dpConvergenceWindow = window
    p.rampSteps = ramp
    p.dpConvergencePercent = pct
    p.usePressureDropSlopeGate = gate
    p.dpSlopeWindowIters = slope_window
    p.dpSlopeMaxDelta = slope_max
    return p


def test_dp_window_constant_drop_stops_when_the_window_fills():
    # iterations 5..24 fill the 20-iteration window
    assert op.dp_exit_iteration(dp_params(), np.full(100, 250.0)) == 24
    assert op.dp_exit_iteration(dp_params(ramp=0), np.full(100, 250.0)) == 19


def test_dp_window_spread_and_latest_deviation():
    # +-2 % alternation: the standard deviation is 2 % of the mean
    alt = 250.0 * (1 + 0.02 * (-1.0) ** np.arange(200))
    assert op.dp_exit_iteration(dp_params(), alt) == -1
    assert op.dp_exit_iteration(dp_params(pct=2.5), alt) == 24
    # a 100-iteration window whose newest value jumps by 5 %: the spread is
    # 0.5 % of the mean, but the newest value is 4.95 % from it, so the exit
    # waits one iteration, until the newest value is back at the mean
    d = np.full(200, 250.0)
    d[104] = 262.5
    assert op.dp_exit_iteration(dp_params(window=100), d) == 105
    assert op.dp_exit_iteration(dp_params(window=100), np.full(200, 250.0)) == 104
    # |mean| below 1 Pa: percentages are of 1 Pa
    tiny = 1e-3 * (1 + 0.5 * (-1.0) ** np.arange(40))
    assert op.dp_exit_iteration(dp_params(), tiny) == 24


def test_dp_slope_gate():
    """A drop drifting by 0.01 Pa per iteration passes the 1 % window test;
    the slope gate (|slope| <= maxDelta / slopeWindow) holds the exit for
    maxDelta = 0.05 (limit 0.0025 per iteration) and lets it pass for 0.5."""
    drift = 250.0 + 0.01 * np.arange(200)
    assert op.dp_exit_iteration(dp_params(), drift) == 24
    assert op.dp_exit_iteration(dp_params(gate=1, slope_max=0.05), drift) == -1
    assert op.dp_exit_iteration(dp_params(gate=1, slope_max=0.5), drift) == 24
