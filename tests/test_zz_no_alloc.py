"""cp_run_noalloc (the solve that allocates no device memory) against the
allocating path it varies, cp_create + cp_set_fields + cp_run +
cp_get_fields: two calls in a row, the second on a smaller grid, each bit
for bit equal to the context's result, and an oversized grid rejected with
CP_E_CAPACITY. The second call reuses the scratch the first one left."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import cpbind as cb
import workloads as wl

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def context_run(p, ct, u0, v0, p0):
    g = cb.Solver(p, ct)
    g.set_fields(u0, v0, p0)
    st = g.run()
    out = (g.field(cb.F_U), g.field(cb.F_V), g.field(cb.F_P))
    g.close()
    return out, st


def noalloc_run(p, ct, u0, v0, p0):
    u, v, pp = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (u0, v0, p0))
    st = cb.StepStats()
    pr = p.copy()
    rc = cb.load().cp_run_noalloc(C.byref(pr), cb.dptr(np.ascontiguousarray(ct, dtype=np.float64)), cb.dptr(u),
                                  cb.dptr(v), cb.dptr(pp), C.byref(st))
    cb.check(rc)
    return (u, v, pp), st.as_dict()


def start_fields(p, seed):
    rng = np.random.default_rng(seed)
    M, N = p.M, p.N
    return (0.01 * rng.random((M + 1, N)), 0.001 * rng.random((M, N + 1)), rng.random((M + 1, N + 1)))


@pytest.mark.parametrize("precision", [cb.CP_PREC_FP64, cb.CP_PREC_MIXED], ids=["fp64", "mixed"])
def test_noalloc_equals_context_twice_second_smaller(precision):
    cases = [wl.network(64, 128, 3, precision), wl.network(40, 96, 7, precision)]
    for k, (p, ct) in enumerate(cases):
        q = wl.fast_mode(p, 3, 4, 1) if precision == cb.CP_PREC_MIXED else p.copy()
        q.maxIterations = 40
        q.epsilon = 1e-12
        u0, v0, p0 = start_fields(q, k)
        say(f"[noalloc] call {k}: {q.M} x {q.N}, precision {precision}")
        ref, st_ref = context_run(q, ct, u0, v0, p0)
        got, st = noalloc_run(q, ct, u0, v0, p0)
        assert st["iter"] == st_ref["iter"] == 40
        for key in ("itU", "itV", "itP", "error", "residMass", "pseudoCFL", "health"):
            assert st[key] == st_ref[key], (k, key, st[key], st_ref[key])
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)


def test_noalloc_rejects_a_grid_beyond_the_scratch():
    p, ct = wl.network(1024, 2048, 1)
    u = np.zeros((p.M + 1, p.N)); v = np.zeros((p.M, p.N + 1)); pp = np.zeros((p.M + 1, p.N + 1))
    st = cb.StepStats()
    rc = cb.load().cp_run_noalloc(C.byref(p), cb.dptr(ct), cb.dptr(u), cb.dptr(v), cb.dptr(pp), C.byref(st))
    assert rc == cb.CP_E_CAPACITY
    assert b"CP_NOALLOC_SCRATCH_BYTES" in cb.load().cp_last_error(None)
