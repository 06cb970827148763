"""Pins of the oracle's aggregation-multigrid preconditioner (CPU).

The V-cycle is checked against what the mathematics fixes, independently of
the oracle's loops:
  * every coarse operator is the Galerkin product P^T A P of the aggregation
    matrix P (2x2 piecewise-constant aggregates);
  * the V-cycle equals the two- and three-grid operators composed from dense
    matrices (damped Jacobi before and after, scaled coarse correction,
    exact solve on the coarsest level);
  * the V-cycle is symmetric positive definite;
  * CG preconditioned with it reaches the dense direct solve of the pressure
    system, and SIMPLE reaches the same fixed point with either
    preconditioner.
"""
from __future__ import annotations

import numpy as np
import pytest

import cpbind as cb
import oracle_py as op
import workloads as wl


def random_operator(H, W, seed):
    """5-point symmetric M-matrix with high-contrast couplings and a
    Dirichlet column on the east side (as the p' system's outlet)."""
    rng = np.random.default_rng(seed)
    cE = 10.0 ** rng.uniform(-4, 0, (H, W))
    cN = 10.0 ** rng.uniform(-4, 0, (H, W))
    cE[:, -1] = 0.0
    cN[-1, :] = 0.0
    aP = np.zeros((H, W))
    aP += cE
    aP[:, 1:] += cE[:, :-1]
    aP += cN
    aP[1:, :] += cN[:-1, :]
    aP[:, -1] += 10.0 ** rng.uniform(-4, 0, H)  # Dirichlet outlet
    return aP, cE, cN


def dense(aP, cE, cN):
    H, W = aP.shape
    n = H * W
    A = np.zeros((n, n))
    for I in range(H):
        for J in range(W):
            k = I * W + J
            A[k, k] = aP[I, J]
            if J + 1 < W:
                A[k, k + 1] = A[k + 1, k] = -cE[I, J]
            if I + 1 < H:
                A[k, k + W] = A[k + W, k] = -cN[I, J]
    return A


def aggregation(H, W):
    Hc, Wc = (H + 1) // 2, (W + 1) // 2
    P = np.zeros((H * W, Hc * Wc))
    for I in range(H):
        for J in range(W):
            P[I * W + J, (I // 2) * Wc + J // 2] = 1.0
    return P, Hc, Wc


def dense_vcycle(A, shapes, r, omega, scale):
    """V-cycle from dense matrices: shapes = [(H0, W0), (H1, W1), ...]."""
    H, W = shapes[0]
    if len(shapes) == 1:
        return np.linalg.solve(A, r)
    Dinv = 1.0 / np.diag(A)
    x = omega * Dinv * r
    x = x + omega * Dinv * (r - A @ x)
    P, _, _ = aggregation(H, W)
    Ac = P.T @ A @ P
    ec = dense_vcycle(Ac, shapes[1:], P.T @ (r - A @ x), omega, scale)
    x = x + scale * (P @ ec)
    x = x + omega * Dinv * (r - A @ x)
    x = x + omega * Dinv * (r - A @ x)
    return x


@pytest.mark.parametrize("H,W", [(9, 13), (16, 24), (7, 30)])
def test_coarse_operators_are_galerkin_products(H, W):
    aP, cE, cN = random_operator(H, W, H * W)
    A = dense(aP, cE, cN)
    level = 0
    while True:
        try:
            cP, cEc, cNc, nl = op.mg_level(aP, cE, cN, level + 1)
        except ValueError:
            break
        h, w = (aP.shape if level == 0 else prev.shape)
        P, Hc, Wc = aggregation(h, w)
        assert cP.shape == (Hc, Wc)
        Ac = P.T @ A @ P
        np.testing.assert_allclose(dense(cP, cEc, cNc), Ac, rtol=1e-12, atol=1e-14 * np.abs(Ac).max())
        A, prev, level = Ac, cP, level + 1
    assert level >= 1


@pytest.mark.parametrize("H,W,omega,scale", [(6, 10, 0.8, 1.8), (12, 20, 0.8, 1.8), (11, 23, 0.67, 1.0)])
def test_vcycle_equals_dense_composition(H, W, omega, scale):
    aP, cE, cN = random_operator(H, W, 7 + H)
    r = np.random.default_rng(H * W).standard_normal((H, W))
    z, nl = op.mg_apply(aP, cE, cN, r, omega, scale)
    shapes = [(H, W)]
    while len(shapes) < nl:
        h, w = shapes[-1]
        shapes.append(((h + 1) // 2, (w + 1) // 2))
    ref = dense_vcycle(dense(aP, cE, cN), shapes, r.ravel(), omega, scale).reshape(H, W)
    np.testing.assert_allclose(z, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_vcycle_is_symmetric_positive_definite():
    aP, cE, cN = random_operator(40, 70, 3)
    rng = np.random.default_rng(5)
    for _ in range(3):
        u, w = rng.standard_normal((2, 40, 70))
        zu, _ = op.mg_apply(aP, cE, cN, u)
        zw, _ = op.mg_apply(aP, cE, cN, w)
        assert abs(np.vdot(w, zu) - np.vdot(u, zw)) <= 1e-11 * abs(np.vdot(u, zu))
        assert np.vdot(u, zu) > 0


def test_vcycle_fp32_tracks_fp64():
    aP, cE, cN = random_operator(33, 65, 11)
    r = np.random.default_rng(2).standard_normal((33, 65))
    z64, _ = op.mg_apply(aP, cE, cN, r)
    z32, _ = op.mg_apply(aP, cE, cN, r, fp32=True)
    assert np.abs(z32 - z64).max() <= 1e-4 * np.abs(z64).max()


def _mg(p):
    q = p.copy()
    q.pressurePrecond = cb.CP_PRECOND_AMG
    return q


def test_mg_pcg_reaches_direct_solve():
    """With a loose iteration cap and a tight tolerance, MG-preconditioned CG
    solves the assembled p' system of a network to the dense solution."""
    p, ct = wl.network(24, 40, 3)
    p = _mg(p)
    p.pTol = 1e-14
    p.maxPressureIter = 200
    o = op.Oracle(p, ct)
    o.run(5)
    o.momentum()
    o.pressure()
    s = o.system(2)
    n = s["n"]
    A = np.zeros((n, n))
    for r in range(n):
        for k in range(s["row_ptr"][r], s["row_ptr"][r + 1]):
            A[r, s["col_idx"][k]] = s["values"][k]
    x = np.linalg.solve(A, s["rhs"])
    np.testing.assert_allclose(s["sol"], x, rtol=0, atol=1e-9 * np.abs(x).max())


def test_mg_reaches_the_same_fixed_point():
    """SIMPLE converged with the MG preconditioner has the fixed point of the
    Jacobi-preconditioned (reference) solver: the preconditioner changes the
    path, not the discrete equations."""
    p, ct = wl.network(20, 48, 5)
    p.massResidualScope = 1
    p.epsilon = 1e-11
    p.pTol = 1e-13
    p.maxPressureIter = 300
    p.maxIterations = 4000
    res = []
    for prec in (cb.CP_PRECOND_JACOBI, cb.CP_PRECOND_AMG):
        q = p.copy()
        q.pressurePrecond = prec
        o = op.Oracle(q, ct)
        n, st, _ = o.run(4000)
        assert st["converged"] == 1, st
        res.append([o.field(f) for f in (cb.F_U, cb.F_V, cb.F_P)])
    for a, b in zip(*res):
        assert np.abs(a - b).max() <= 1e-7 * np.abs(a).max()
