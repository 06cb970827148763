"""Pins of the thermal oracle (oracle/thermal_oracle.hpp): the reference's
conjugate-heat-transfer system (ThermalSolver/thermal_solver.cpp:421-604)
checked against what the physics and the reference's own closed forms fix,
not against a retyped formula:

* conductivity blends: the linear column of the table in the reference's
  comment (thermal_solver.cpp:250-258), the RAMP end points, the harmonic
  mean at gamma = 1/2, and the RAMP slope against the reference's own
  derivative dk_dgamma (:316-320). The RAMP columns of that table do not
  follow the formula printed next to them (gamma = 0.75, q = 1: 66.8 in the
  table, 57.7 from the formula), so only its linear column is used;
* the lateral conductance between unlike cells: two half cells in series;
* structure: an M-matrix sign pattern, a symmetric matrix without flow;
* conservation: every column of A sums to zero except where heat leaves
  the domain (outlet advection, inlet face), so the heat the base takes in
  equals what leaves; with the direct solution, the reference's outlet heat
  Q_out plus the heat conducted out through the inlet face equals Q_in;
* a developed channel: far from both ends the temperature rises by
  q0 dx / (rho Cp U sum_k po_k dz_k) per cell, the heat balance of one slice.

AMGCL (the reference's solver) is absent, and the solution of the system
does not depend on the solver: the checks solve it directly (scipy)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

ROOT = Path(__file__).resolve().parents[1]
sys.path[:0] = [str(ROOT / "oracle"), str(ROOT / "coldplate-solver_amd")]
import oracle_py as op  # noqa: E402


def params(Nx, Ny, **kw):
    p = op.thermal_default_params()
    p.Nx, p.Ny = Nx, Ny
    p.dx = p.dy = 1e-4
    p.q0 = 1e5
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def dz_of(p):
    return np.array([p.H_b / p.nz_solid] * p.nz_solid + [p.H_t / p.nz_fluid] * p.nz_fluid)


def pois_of(p):
    k = np.arange(p.nz_fluid)
    z = (k + 0.5) / p.nz_fluid
    return np.concatenate([np.zeros(p.nz_solid), 4 * z * (1 - z)])


def conductivity(g, mode, qk=1.0):
    """kval(gamma) read back from the assembled lateral coupling of two equal
    fluid cells without flow: A[P, W] = -harm(k, k) dy dz / dx = -k dy dz / dx."""
    p = params(2, 1, nz_solid=1, nz_fluid=1, k_interp_mode=mode, qk=qk)
    gam = np.full((1, 2), g)
    A, _ = op.thermal_assemble(p, gam, np.zeros((1, 2)), np.zeros((1, 2)))
    P, W = 1 * 2 + 1, 1 * 2 + 0  # cell (1, 0, 1) and its west neighbour
    return -A[P, W] * p.dx / (p.dy * p.H_t)


def test_linear_blend_matches_reference_table():
    for g, want in ((0.0, 400.0), (0.25, 300.2), (0.5, 200.3), (0.75, 100.5), (1.0, 0.6)):
        assert conductivity(g, 0) == pytest.approx(want, abs=0.05 + 1e-9)  # one decimal


def test_ramp_end_points_and_slope():
    assert conductivity(0.0, 1) == pytest.approx(400.0, rel=1e-12)
    assert conductivity(1.0, 1) == pytest.approx(0.6, rel=1e-12)
    for qk in (1.0, 10.0, 30.0):
        for g in (0.2, 0.5, 0.8):
            h = 1e-6
            slope = (conductivity(g + h, 1, qk) - conductivity(g - h, 1, qk)) / (2 * h)
            # dk_dgamma, thermal_solver.cpp:316-320
            want = (400.0 - 0.6) * (-(1.0 + qk)) / (1.0 + qk * g) ** 2
            assert slope == pytest.approx(want, rel=1e-6)
    # RAMP is convex in gamma: below the linear blend inside (0, 1)
    for g in (0.25, 0.5, 0.75):
        assert conductivity(g, 1) < conductivity(g, 0)


def test_harmonic_blend_is_the_harmonic_mean_at_one_half():
    assert conductivity(0.5, 2) == pytest.approx(2 * 400 * 0.6 / 400.6, rel=1e-12)


def test_lateral_coupling_is_two_half_cells_in_series():
    """Flux continuity across a face between a solid and a fluid cell: the
    conductance is that of the two half cells in series,
    1 / (dx / (2 k1 A) + dx / (2 k2 A))."""
    p = params(2, 1, nz_solid=1, nz_fluid=1, k_interp_mode=0)
    A, _ = op.thermal_assemble(p, np.array([[0.0, 1.0]]), np.zeros((1, 2)), np.zeros((1, 2)))
    area = p.dy * p.H_t
    want = 1.0 / (p.dx / (2 * p.k_s * area) + p.dx / (2 * p.k_f * area))
    assert -A[3, 2] == pytest.approx(want, rel=1e-12)
    assert -A[2, 3] == pytest.approx(want, rel=1e-12)


def channel_fields(Nx, Ny, seed=0, buffers=True):
    """Fins across a channel: gamma = 1 (fluid) with solid posts, a few
    intermediate (buffer) values, and a velocity field that is zero in
    solid cells."""
    rng = np.random.default_rng(seed)
    gam = np.ones((Ny, Nx))
    for _ in range(max(1, Nx * Ny // 60)):
        j, i = rng.integers(1, Ny - 1), rng.integers(2, Nx - 2)
        gam[j - 1:j + 1, i - 1:i + 1] = 0.0
    if buffers:
        m = rng.random((Ny, Nx)) < 0.05
        gam[m] = rng.uniform(0.0, 1.0, m.sum())
    u = np.where(gam > 0.01, 0.05 + 0.02 * rng.random((Ny, Nx)), 0.0)
    v = np.where(gam > 0.01, 0.01 * rng.standard_normal((Ny, Nx)), 0.0)
    return gam, u, v


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_m_matrix_sign_pattern(mode):
    p = params(17, 9, k_interp_mode=mode)
    g, u, v = channel_fields(17, 9)
    A, b = op.thermal_assemble(p, g, u, v)
    A = A.tocoo()
    off = A.row != A.col
    assert (A.data[off] <= 0).all()
    assert (A.diagonal() > 0).all()
    assert np.bincount(A.row, minlength=A.shape[0]).max() <= 7
    assert (b >= 0).all()


def test_symmetric_without_flow():
    p = params(13, 7)
    g, _, _ = channel_fields(13, 7)
    A, _ = op.thermal_assemble(p, g, np.zeros_like(g), np.zeros_like(g))
    assert abs(A - A.T).max() <= 1e-15 * abs(A).max()


def boundary_outflow(p, g, u, v):
    """Conductance through which heat leaves each cell of the domain: the
    outlet's advected heat rho Cp u A, the inlet face's conductance and its
    reverse flow; zero elsewhere (the walls and the top are adiabatic)."""
    assert p.k_interp_mode == 0  # the inlet conductance below is the linear blend's
    Nz = p.nz_solid + p.nz_fluid
    dz, po = dz_of(p), pois_of(p)
    out = np.zeros((Nz, p.Ny, p.Nx))
    rc = p.rho * p.Cp
    fl = g > 0.01
    gc = np.clip(g, 0, 1)
    kf = p.k_s * (1 - gc) + p.k_f * gc
    for k in range(p.nz_solid, Nz):
        Ax, Ay = p.dy * dz[k], p.dx * dz[k]
        uP = np.where(gc > 0.01, u * po[k], 0.0)
        vP = np.where(gc > 0.01, v * po[k], 0.0)
        out[k, :, -1] += np.where(fl[:, -1] & (uP[:, -1] > 0), rc * uP[:, -1] * Ax, 0.0)
        out[k, -1, :] += np.where(fl[-1, :] & (vP[-1, :] > 0), rc * vP[-1, :] * Ay, 0.0)
        Dbc = 2.0 * kf[:, 0] * Ax / p.dx
        out[k, :, 0] += np.where(fl[:, 0], Dbc + np.maximum(-rc * uP[:, 0] * Ax, 0.0), 0.0)
    return out.ravel()


def test_columns_sum_to_the_boundary_outflow():
    """Interior fluxes cancel in pairs (each face's conduction and upwinded
    advection leave one cell and enter the next), so a column of A sums to
    the conductance through which heat leaves the domain at that cell."""
    p = params(23, 11, k_interp_mode=0)
    g, u, v = channel_fields(23, 11)
    A, _ = op.thermal_assemble(p, g, u, v)
    cols = np.asarray(A.sum(axis=0)).ravel() - 1e-12
    want = boundary_outflow(p, g, u, v)
    scale = abs(A.diagonal()).max()
    assert np.abs(cols - want).max() <= 1e-12 * scale


def test_heat_balance_of_the_direct_solution():
    """Q_in (the base) = Q_out (outlet advection, the reference's formula)
    + the heat conducted out through the inlet face, for a flow that
    conserves mass (straight channels between solid and buffer fins)."""
    p = params(40, 12, k_interp_mode=0)
    g = np.ones((12, 40))
    g[3:5] = 0.0   # solid fin rows
    g[8] = 0.5     # a buffer row: fluid with half the conductivity blend
    rng = np.random.default_rng(3)
    u = np.where(g > 0.01, (0.02 + 0.03 * rng.random(12))[:, None], 0.0) * np.ones((1, 40))
    v = np.zeros_like(u)
    A, b = op.thermal_assemble(p, g, u, v)
    T = spla.spsolve(A.tocsc(), b)
    st = op.thermal_stats(p, g, u, v, T)
    Tc = T.reshape(p.nz_solid + p.nz_fluid, p.Ny, p.Nx)
    dz = dz_of(p)
    gc = np.clip(g[:, 0], 0, 1)
    kf = p.k_s * (1 - gc) + p.k_f * gc
    q_inlet = 0.0
    for k in range(p.nz_solid, len(dz)):
        Dbc = 2.0 * kf * p.dy * dz[k] / p.dx
        q_inlet += np.sum(np.where(g[:, 0] > 0.01, Dbc * (Tc[k, :, 0] - p.T_inlet), 0.0))
    assert st["Q_in"] == pytest.approx(p.q0 * p.Nx * p.Ny * p.dx * p.dy, rel=1e-14)
    # the 1e-12 diagonal shift of every row takes 1e-12 sum(T) out as well
    leak = 1e-12 * T.sum()
    assert st["Q_out"] + q_inlet + leak == pytest.approx(st["Q_in"], rel=1e-9)
    assert 0 < q_inlet < st["Q_in"]
    assert st["imbalance_pct"] == pytest.approx(100 * (st["Q_in"] - st["Q_out"]) / st["Q_in"], rel=1e-12)


def test_developed_channel_warms_at_the_slice_heat_balance_rate():
    """All fluid, uniform U: far from the inlet and the outlet every layer
    warms by q0 dx / (rho Cp U sum_k po_k dz_k) per cell."""
    U = 0.01
    # wide rows (dy = 1 cm) make the heat the diagonal shift takes out of a
    # cell (1e-12 T) negligible beside what the base puts in (q0 dx dy)
    p = params(640, 3, k_interp_mode=0, k_s=4.0, H_t=2e-4, nz_fluid=5, dy=1e-2)
    g = np.ones((3, 640))
    u = np.full((3, 640), U)
    A, b = op.thermal_assemble(p, g, u, np.zeros_like(g))
    T = spla.spsolve(A.tocsc(), b).reshape(p.nz_solid + p.nz_fluid, 3, 640)
    carried = p.rho * p.Cp * U * p.dy * np.sum(pois_of(p) * dz_of(p))
    d = np.diff(T[:, 1, 99:300], axis=1)  # the inlet's and outlet's reach: < 100 cells
    assert np.abs(d / (p.q0 * p.dx * p.dy / carried) - 1).max() < 1e-7


def test_stats_of_closed_form_fields():
    p = params(8, 4, k_interp_mode=0)
    g = np.ones((4, 8))
    g[:, 3] = 0.0  # one solid column: fins
    u = np.where(g > 0.01, 0.02, 0.0)
    Nz = p.nz_solid + p.nz_fluid
    T = np.full((Nz, 4, 8), p.T_inlet + 5.0)
    st = op.thermal_stats(p, g, u, np.zeros_like(u), T)
    mdot = p.rho * 0.02 * 4 * p.dy * np.sum(pois_of(p) * dz_of(p))
    assert st["T_out"] == pytest.approx(p.T_inlet + 5.0, rel=1e-14)
    assert st["Q_out"] == pytest.approx(mdot * p.Cp * 5.0, rel=1e-12)
    # a field linear in the layer index: volume averages by region
    T = np.broadcast_to(np.arange(Nz, dtype=float)[:, None, None], (Nz, 4, 8)).copy()
    st = op.thermal_stats(p, g, u, np.zeros_like(u), T)
    dz = dz_of(p)
    assert st["T_avg_base"] == pytest.approx(np.mean(np.arange(p.nz_solid)), rel=1e-14)
    fl = np.arange(p.nz_solid, Nz)
    assert st["T_avg_fluid"] == pytest.approx(np.sum(fl * dz[fl]) / np.sum(dz[fl]), rel=1e-14)
    assert st["T_avg_solid"] == pytest.approx(st["T_avg_fluid"], rel=1e-14)
    assert st["T_avg_global"] == pytest.approx(np.sum(np.arange(Nz) * dz) / dz.sum(), rel=1e-14)
    assert (st["T_base_min"], st["T_base_max"]) == (0.0, p.nz_solid - 1.0)
    assert (st["T_fin_min"], st["T_fin_max"]) == (float(p.nz_solid), Nz - 1.0)
