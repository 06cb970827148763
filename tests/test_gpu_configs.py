"""GPU tests at the sizes of the batched, served and 2048^2 configurations,
through properties that hold at any size (the CPU oracle is pinned against
the device on smaller grids in the other test files):
  * 2048 x 2048 gamma field: from the same state, one mixed-precision outer
    iteration changes the fields as one FP64 outer iteration does, and the
    boundary conditions hold;
  * batched mix of networks and gamma fields of log-uniform sizes: every
    slot converges, and a slot's result is bit-identical alone;
  * served stream with Poisson arrivals over straight / network / gamma
    requests: every result is bit-identical to the same request alone.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import cpbind as cb
import workloads as wl

pytestmark = pytest.mark.gpu


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


def test_batched_mix_converges_and_equals_alone():
    geoms = sorted(wl.batch_64(), key=lambda g: g[0].M * g[0].N)[:12]
    reqs = [(served(p, 1e-5, 6000), ct, 0.0, 0.5) for p, ct in geoms]
    say("[configs] 12 smallest batch-64 geometries on 6 slots")
    b = cb.Batch(cb.CP_PREC_MIXED, 1 << 30, 6)
    res = b.serve(reqs, 20)
    b.close()
    st = [(r["status"], r["iter"]) for r in res]
    say(f"  status/iterations {st}")
    assert all(r["status"] in (cb.SLOT_CONVERGED, cb.SLOT_MAXITER) and r["rc"] == cb.CP_OK for r in res), st
    assert sum(r["status"] == cb.SLOT_CONVERGED for r in res) >= 9, st
    for k in (0, 7):
        b = cb.Batch(cb.CP_PREC_MIXED, 256 << 20, 1)
        alone = b.serve([reqs[k]], 20)[0]
        b.close()
        assert (alone["iter"], alone["dp"]) == (res[k]["iter"], res[k]["dp"])


def test_served_poisson_stream_equals_alone():
    specs = wl.stream_2000(10, 7, max_cells=96 * 256)
    arr = wl.poisson_arrivals(10, 40.0, 7)
    reqs = []
    for r, t in zip(specs, arr):
        p, ct = wl.build_request(r)
        reqs.append((served(p, 1e-5, 3000), ct, float(t), 0.5 if r["kind"] == 2 else 1.0 - 1e-6))
    say("[configs] Poisson stream of 10 requests on 3 slots")
    b = cb.Batch(cb.CP_PREC_MIXED, 512 << 20, 3)
    mix = b.serve(reqs, 20)
    b.close()
    for k, (r, req) in enumerate(zip(mix, reqs)):
        assert r["rc"] == cb.CP_OK and r["tAdmit"] >= req[2] - 1e-9
        b = cb.Batch(cb.CP_PREC_MIXED, 128 << 20, 1)
        alone = b.serve([(req[0], req[1], 0.0, req[3])], 20)[0]
        b.close()
        assert (alone["status"], alone["iter"], alone["dp"]) == (r["status"], r["iter"], r["dp"]), k
