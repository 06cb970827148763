"""GPU tests of the request loop (cp_serve): continuous batching with
admission by HBM footprint. Every result served in the mix equals the same
request served alone, bit for bit (per-slot reductions depend only on the
slot's own shape), and requests admitted while others run overlap with them
on the device."""
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


def fast(p):
    q = wl.fast_mode(p, 4, 4, 2)
    q.pressurePrecond = cb.CP_PRECOND_AMG
    q.epsilon = 1e-6
    q.maxIterations = 5000
    return q


def requests():
    out = []
    p, ct = wl.network(64, 128, 31, cb.CP_PREC_MIXED)
    out.append((fast(p), ct, 0.0, 1.0 - 1e-6))
    p, ct = wl.gamma_field(40, 96, 9, cb.CP_PREC_MIXED, corr=5.0)
    out.append((fast(p), ct, 0.0, 0.5))
    p, ct = wl.straight_200x50(cb.CP_PREC_MIXED)
    out.append((fast(p), ct, 0.02, 1.0 - 1e-6))
    p, ct = wl.network(48, 96, 7, cb.CP_PREC_MIXED)
    out.append((fast(p), ct, 0.05, 1.0 - 1e-6))
    p, ct = wl.network(32, 160, 3, cb.CP_PREC_MIXED, U=0.1)
    out.append((fast(p), ct, 0.05, 1.0 - 1e-6))
    return out


def test_served_results_equal_alone():
    reqs = requests()
    say("[serve] mix of 5 requests on 2 slots")
    b = cb.Batch(cb.CP_PREC_MIXED, 256 << 20, 2)
    mix = b.serve(reqs, steps_per_poll=10)
    b.close()
    for k, r in enumerate(mix):
        assert r["rc"] == cb.CP_OK and r["status"] == cb.SLOT_CONVERGED, (k, r)
        assert r["tAdmit"] >= reqs[k][2] - 1e-9
    # continuous batching: some request was admitted while another was running
    overlap = any(a["tAdmit"] > c["tAdmit"] and a["tAdmit"] < c["tDone"]
                  for a in mix for c in mix if a is not c)
    assert overlap
    for k, req in enumerate(reqs):
        say(f"[serve] request {k} alone")
        b = cb.Batch(cb.CP_PREC_MIXED, 64 << 20, 1)
        alone = b.serve([(req[0], req[1], 0.0, req[3])], steps_per_poll=10)[0]
        b.close()
        assert alone["iter"] == mix[k]["iter"], (k, alone, mix[k])
        assert alone["dp"] == mix[k]["dp"], (k, alone["dp"], mix[k]["dp"])


def test_admission_waits_for_footprint_and_rejects_oversized():
    p, ct = wl.network(64, 128, 5, cb.CP_PREC_MIXED)
    q = fast(p)
    big, bct = wl.network(256, 512, 1, cb.CP_PREC_MIXED)
    b0 = cb.Batch(cb.CP_PREC_MIXED, 1 << 20, 1)
    fp = b0.footprint(q.M, q.N)
    b0.close()
    # room for exactly one small geometry at a time, never for the big one
    b = cb.Batch(cb.CP_PREC_MIXED, fp + (1 << 20), 4)
    say("[serve] footprint admission")
    res = b.serve([(q, ct, 0.0, 1.0 - 1e-6), (q, ct, 0.0, 1.0 - 1e-6), (fast(big), bct, 0.0, 1.0 - 1e-6)],
                  steps_per_poll=10)
    b.close()
    a, c, d = res
    assert a["rc"] == cb.CP_OK and c["rc"] == cb.CP_OK
    assert c["tAdmit"] >= a["tDone"] - 1e-9  # second waited for the first's space
    assert a["dp"] == c["dp"] and a["iter"] == c["iter"]
    assert d["rc"] == cb.CP_E_CAPACITY and d["slot"] == -1


def test_pool_results_equal_alone_and_streams_overlap():
    """cp_serve_pool: the 5 requests over two 1-slot batches (two host
    threads, two streams). Each result equals the request served alone, bit
    for bit, and the two batches ran at the same time."""
    reqs = requests()
    say("[serve] pool of 2 batches x 1 slot, 5 requests")
    bs = [cb.Batch(cb.CP_PREC_MIXED, 128 << 20, 1) for _ in range(2)]
    for b in bs:
        # timing bursts of both probed kernels take the GPU alone and rotate
        # over both batches; results unchanged
        b.probe((cb.K_PCG_Q, cb.K_MOM_ASM))
    mix = cb.serve_pool(bs, reqs, steps_per_poll=10)
    recs = bs[0].probe_read()
    bursts = [len(b.probe_read()) for b in bs]
    assert {r["kernel"] for r in recs} == {cb.K_PCG_Q, cb.K_MOM_ASM}
    assert all(r["ms"] > 0 and r["units"] > 0 and r["launches"] >= 16 for r in recs)
    # a round launches once per batch with running slots: both, while both run
    assert max(r["launches"] for r in recs) == 32
    for b in bs:
        b.close()
    for k, r in enumerate(mix):
        assert r["rc"] == cb.CP_OK and r["status"] == cb.SLOT_CONVERGED, (k, r)
        assert r["tAdmit"] >= reqs[k][2] - 1e-9
    assert sum(bursts) > 0
    # each batch has one slot, so overlapping runs were on different streams
    overlap = any(a["tAdmit"] < c["tDone"] and c["tAdmit"] < a["tDone"]
                  for a in mix for c in mix if a is not c)
    assert overlap
    for k, req in enumerate(reqs):
        b = cb.Batch(cb.CP_PREC_MIXED, 64 << 20, 1)
        alone = b.serve([(req[0], req[1], 0.0, req[3])], steps_per_poll=10)[0]
        b.close()
        assert alone["iter"] == mix[k]["iter"], (k, alone, mix[k])
        assert alone["dp"] == mix[k]["dp"], (k, alone["dp"], mix[k]["dp"])


def test_pool_rejects_a_busy_batch():
    p, ct = wl.network(48, 96, 7, cb.CP_PREC_MIXED)
    a = cb.Batch(cb.CP_PREC_MIXED, 64 << 20, 1)
    b = cb.Batch(cb.CP_PREC_MIXED, 64 << 20, 1)
    b.admit(fast(p), ct)
    with pytest.raises(cb.CpError):
        cb.serve_pool([a, b], [(fast(p), ct, 0.0, 1.0 - 1e-6)])
    a.close()
    b.close()


def test_pool_rejects_an_invalid_request_and_serves_the_rest():
    """An invalid request (uvAlpha = 0) in the shared queue gets its error code
    and no slot; the other requests, arriving over time, all converge."""
    reqs = requests()[:4]
    bad = reqs[1][0].copy()
    bad.uvAlpha = 0.0
    reqs.insert(2, (bad, reqs[1][1], 0.01, 1.0 - 1e-6))
    say("[serve] pool with an invalid request")
    bs = [cb.Batch(cb.CP_PREC_MIXED, 128 << 20, 1) for _ in range(2)]
    res = cb.serve_pool(bs, reqs, steps_per_poll=10)
    for b in bs:
        b.close()
    assert res[2]["rc"] == cb.CP_E_ARG and res[2]["slot"] == -1 and res[2]["status"] == cb.SLOT_FAILED
    for k in (0, 1, 3, 4):
        assert res[k]["rc"] == cb.CP_OK and res[k]["status"] == cb.SLOT_CONVERGED, (k, res[k])
