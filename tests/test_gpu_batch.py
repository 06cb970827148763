"""GPU tests of the batched form: packing, graph replay, identity alone vs
mixed, admission by footprint, eviction."""
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


def geoms():
    a = wl.network(64, 128, 21, cb.CP_PREC_MIXED)
    b = wl.gamma_field(40, 96, 5, cb.CP_PREC_MIXED, corr=5.0)
    c = wl.straight_200x50(cb.CP_PREC_MIXED)
    return [a, b, c]


@pytest.mark.parametrize("precision", [cb.CP_PREC_FP64, cb.CP_PREC_MIXED], ids=["fp64", "mixed"])
def test_alone_equals_mixed_bitwise(precision):
    """The same geometry gives bit-identical fields alone and packed with
    others (per-slot reduction order depends only on the slot's shape)."""
    gs = []
    for p, ct in geoms():
        q = wl.fast_mode(p) if precision == cb.CP_PREC_MIXED else p.copy()
        q.precision = precision
        gs.append((q, ct))
    say(f"[batch] alone vs mixed, precision {precision}")
    alone = cb.Batch(precision, 256 << 20, 4)
    s0 = alone.admit(*gs[0])
    alone.run(25)
    ua, va, pa = alone.fields(s0)
    mixed = cb.Batch(precision, 256 << 20, 4)
    order = [mixed.admit(*gs[2]), mixed.admit(*gs[0]), mixed.admit(*gs[1])]
    mixed.run(25)
    ub, vb, pb = mixed.fields(order[1])
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(va, vb)
    np.testing.assert_array_equal(pa, pb)
    info = mixed.poll()
    assert info[order[1]]["iter"] == 25


def test_graph_batch_matches_polled_drop_in():
    """FP64, reference budgets: 20 graph-replayed outer iterations of a
    one-slot batch equal 20 polled cp_step calls."""
    p, ct = wl.straight_200x50()
    say("[batch] graph vs polled FP64")
    b = cb.Batch(cb.CP_PREC_FP64, 64 << 20, 1)
    b.set_tiling(1)  # the single-geometry context of a 200 x 50 grid uses R = 1
    s = b.admit(p, ct)
    b.run(20)
    g = cb.Solver(p, ct)
    for _ in range(20):
        g.step()
    u, v, pp = b.fields(s)
    # same kernels in the same order (the graph's extra launches are no-ops)
    np.testing.assert_array_equal(u, g.field(cb.F_U))
    np.testing.assert_array_equal(pp, g.field(cb.F_P))


def test_batch_fp64_matches_oracle():
    p, ct = wl.network(64, 128, 3)
    b = cb.Batch(cb.CP_PREC_FP64, 64 << 20, 2)
    s = b.admit(p, ct)
    b.run(12)
    o = op.Oracle(p, ct)
    o.run(12)
    u, v, pp = b.fields(s)
    for a, f in ((u, cb.F_U), (v, cb.F_V), (pp, cb.F_P)):
        ref = o.field(f)
        assert np.abs(a - ref).max() <= 1e-8 * np.abs(ref).max()


def test_admission_by_footprint_and_evict():
    p, ct = wl.network(64, 128, 9, cb.CP_PREC_MIXED)
    q = wl.fast_mode(p)
    b = cb.Batch(cb.CP_PREC_MIXED, 64 << 20, 8)
    fp = b.footprint(q.M, q.N)
    cap = (64 << 20) // fp
    slots = []
    while True:
        s = b.admit(q, ct)
        if s is None:
            break
        slots.append(s)
    assert 1 <= len(slots) <= min(8, cap)
    b.run(3)
    info = b.poll()
    assert all(info[s]["status"] == cb.SLOT_RUNNING for s in slots)
    b.evict(slots[0])
    assert b.admit(q, ct) == slots[0]
    b.run(2)
    info = b.poll()
    assert info[slots[0]]["iter"] == 2 and info[slots[1]]["iter"] == 5


def test_converged_slot_is_frozen():
    """A converged slot is skipped by later launches: its fields stop
    changing while other slots keep running."""
    p, ct = wl.straight_200x50(cb.CP_PREC_MIXED)
    q = wl.fast_mode(p)
    q.epsilon = 1e-3
    r = wl.fast_mode(p)
    r.epsilon = 1e-12
    b = cb.Batch(cb.CP_PREC_MIXED, 128 << 20, 2)
    s0, s1 = b.admit(q, ct), b.admit(r, ct)
    for _ in range(40):
        b.run(5)
        if b.poll()[s0]["status"] == cb.SLOT_CONVERGED:
            break
    info = b.poll()
    assert info[s0]["status"] == cb.SLOT_CONVERGED
    it0 = info[s0]["iter"]
    u0 = b.fields(s0)[0]
    b.run(10)
    info = b.poll()
    assert info[s0]["iter"] == it0 and info[s1]["iter"] > it0
    np.testing.assert_array_equal(b.fields(s0)[0], u0)


@pytest.mark.parametrize("rows", [1, 2, 4, 8])
def test_tile_heights_agree_with_oracle(rows):
    """Every tile height (64 x 4R tiles; the stencil kernels march R rows per
    wave with lane shuffles for east/west neighbours) gives the oracle's
    result. A grid that is not a multiple of the tile covers the edges."""
    p, ct = wl.network(70, 150, 13)
    say(f"[batch] tile rows {rows}")
    b = cb.Batch(cb.CP_PREC_FP64, 64 << 20, 1)
    b.set_tiling(rows)
    s = b.admit(p, ct)
    b.run(12)
    o = op.Oracle(p, ct)
    o.run(12)
    u, v, pp = b.fields(s)
    for a, f in ((u, cb.F_U), (v, cb.F_V), (pp, cb.F_P)):
        ref = o.field(f)
        assert np.abs(a - ref).max() <= 1e-8 * np.abs(ref).max(), f


@pytest.mark.parametrize("precision", [cb.CP_PREC_FP64, cb.CP_PREC_MIXED], ids=["fp64", "mixed"])
def test_tile_orders_are_bitwise_identical(precision):
    """Every tile-order mask (XCD-split or launch order, for the node kernels
    and for the p' and multigrid kernels) gives bit-identical fields and inner
    counts: a tile's work and its partials do not depend on which workgroup
    takes it."""
    geoms = [wl.network(64, 128, 5, precision), wl.network(48, 200, 9, precision)]
    say(f"[batch] tile orders, precision {precision}")
    runs = []
    for order in range(8):
        b = cb.Batch(precision, 128 << 20, 2)
        b.set_tile_order(order)
        sl = []
        for p, ct in geoms:
            q = wl.fast_mode(p, 3, 4, 1) if precision == cb.CP_PREC_MIXED else p
            if precision == cb.CP_PREC_MIXED:
                q.pressurePrecond = cb.CP_PRECOND_AMG
            sl.append(b.admit(q, ct))
        b.run(15)
        info = b.poll()
        runs.append([(b.fields(s), info[s]["itU"], info[s]["itV"], info[s]["itP"], info[s]["iter"]) for s in sl])
        b.close()
    for other in runs[1:]:
        for (f0, *c0), (f1, *c1) in zip(runs[0], other):
            assert c0 == c1
            for x, y in zip(f0, f1):
                np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("kernel", [cb.K_PCG_Q, cb.K_MOM_ASM, cb.K_MG_UP], ids=["pcg_q", "mom_assemble", "mg_up"])
def test_timing_burst_leaves_results_unchanged(kernel):
    """Timing bursts of k_pcg_q, k_mom_assemble or the finest level's k_mg_up
    between replays (as bench.py runs them) do not change any slot's
    results."""
    p, ct = wl.network(64, 128, 17, cb.CP_PREC_MIXED)
    q = wl.fast_mode(p)
    q.pressurePrecond = cb.CP_PRECOND_AMG
    say(f"[batch] timing burst of kernel {kernel}")
    runs = []
    for burst in (False, True):
        b = cb.Batch(cb.CP_PREC_MIXED, 64 << 20, 2)
        s = b.admit(q, ct)
        for _ in range(4):
            b.run(5)
            if burst:
                assert b.probe_burst(4, kernel) > 0.0
        runs.append(b.fields(s))
        b.close()
    for x, y in zip(*runs):
        np.testing.assert_array_equal(x, y)


def test_autotune_keeps_results():
    """The autotuner picks a valid tiling; the solve then converges to the
    same pressure drop as with the default tiling (tilings change only the
    summation order of reductions)."""
    p, ct = wl.network(64, 160, 27, cb.CP_PREC_MIXED)
    q = wl.fast_mode(p, 3, 4, 1)
    q.pressurePrecond = cb.CP_PRECOND_AMG
    q.epsilon = 1e-9
    q.pTol = 1e-12
    q.maxIterations = 20000
    say("[batch] autotune")
    dps = []
    for tune in (False, True):
        b = cb.Batch(cb.CP_PREC_MIXED, 128 << 20, 2)
        s = b.admit(q, ct)
        if tune:
            t = b.autotune(1)
            assert t["tile_rows"] in (1, 2, 4, 8) and t["mg_rows"] in (2, 4, 8) and t["wg_per_cu"] in (4, 8, 16)
            assert 0 <= t["tile_order"] <= 7
            assert t["ms_per_step"] > 0
        for _ in range(400):
            b.run(50)
            if b.poll()[s]["status"] != cb.SLOT_RUNNING:
                break
        assert b.poll()[s]["status"] == cb.SLOT_CONVERGED
        dps.append(b.full_dp(s))
        b.close()
    assert abs(dps[1] / dps[0] - 1) <= 1e-7, dps


@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_small_sweep_budgets_match_oracle(sweeps):
    """Graph-replayed FP64 batch with a Jacobi budget of 1, 2 or 3 sweeps:
    the first sweep is made by the momentum assembly and the budget's last
    sweep runs without its own finalisation (the momentum finish takes the
    budget as the count). Step by step, the sweep counts equal the oracle's
    and the fields stay within 1e-8 of it."""
    p, ct = wl.network(64, 128, 3)
    p.maxMomentumIter = sweeps
    say(f"[batch] FP64 graph, {sweeps} Jacobi sweeps")
    b = cb.Batch(cb.CP_PREC_FP64, 64 << 20, 1)
    s = b.admit(p, ct)
    o = op.Oracle(p, ct)
    for step in range(6):
        b.run(1)
        so = o.step()
        info = b.poll()[s]
        assert (info["itU"], info["itV"]) == (so["itU"], so["itV"]), (step, info, so)
    u, v, pp = b.fields(s)
    for a, f in ((u, cb.F_U), (v, cb.F_V), (pp, cb.F_P)):
        ref = o.field(f)
        assert np.abs(a - ref).max() <= 1e-8 * np.abs(ref).max()
    b.close()


@pytest.mark.parametrize(
    "cycles,precond,sweeps,mom,pres,pcg,ptol",
    [(1, cb.CP_PRECOND_AMG, 3, 1, 1, 4, None), (2, cb.CP_PRECOND_AMG, 3, 1, 1, 4, None),
     (2, cb.CP_PRECOND_JACOBI, 3, 1, 1, 4, None), (1, cb.CP_PRECOND_JACOBI, 3, 1, 1, 4, None),
     (1, cb.CP_PRECOND_AMG, 1, 1, 1, 4, None), (1, cb.CP_PRECOND_AMG, 2, 1, 1, 4, None),
     (1, cb.CP_PRECOND_AMG, 3, 0, 1, 4, None), (1, cb.CP_PRECOND_AMG, 3, 1, 0, 4, None),
     (1, cb.CP_PRECOND_AMG, 3, 1, 1, 2, None), (1, cb.CP_PRECOND_JACOBI, 3, 1, 1, 3, None),
     (1, cb.CP_PRECOND_AMG, 3, 1, 1, 5, 0.02), (1, cb.CP_PRECOND_AMG, 3, 1, 1, 5, 0.012)],
    ids=["amg-1", "amg-2", "jacobi-2", "jacobi-1", "amg-1-sweeps1", "amg-1-sweeps2", "explicit-cg", "implicit-sor",
         "amg-1-cg2", "jacobi-1-cg3", "amg-1-cg5-stops", "amg-1-cg5-stop1"])
def test_graph_batch_matches_polled_mixed(cycles, precond, sweeps, mom, pres, pcg, ptol):
    """Mixed precision, `pcg` CG iterations per refinement cycle: the graph
    step equals the polled context bit for bit and with the same inner
    counts. The graph step merges finalisations where a later one can carry
    them, forms u* and v* of an implicit-momentum slot in the pressure
    assembly and the correction instead of storing them, and with one
    refinement cycle reads p' as the CG iterate itself, which it forms late
    (k_pcg_q adds two updates at every even iteration from the third, the
    correction the one or two still pending); the context finalises every
    phase on its own, updates x every iteration and stores u*, v* and p'.
    The cases cover both preconditioners (each its own form of the CG
    stencil), one and two refinement cycles, momentum budgets ending after
    the assembly's sweep, on an even and on an odd sweep, the explicit-
    momentum and SOR backends, which the graph step leaves unfolded, CG
    budgets of 2, 3, 4 and 5 iterations, and a loose CG tolerance under
    which solves stop inside the budget (after 1 and after 3 counted
    iterations: two updates pending, none or two of them added by k_pcg_q)."""
    p, ct = wl.network(64, 128, 3, cb.CP_PREC_MIXED)
    q = wl.fast_mode(p, sweeps, pcg, cycles)
    q.pressurePrecond = precond
    q.momentumSolverType, q.pressureSolverType = mom, pres
    if ptol is not None:
        q.pTol = ptol
    if pres == 0:
        q.maxPressureIter = 20
    say(f"[batch] mixed, {cycles} cycles, precond {precond}, {sweeps} sweeps, momentum {mom}, pressure {pres}, "
        f"{pcg} CG iterations, pTol {q.pTol}: graph vs polled")
    b = cb.Batch(cb.CP_PREC_MIXED, 64 << 20, 1)
    b.set_tiling(1)  # the context's tiling for a 64 x 128 grid
    s = b.admit(q, ct)
    g = cb.Solver(q, ct)
    its = []
    for step in range(15):
        b.run(1)
        sg = g.step()
        info = b.poll()[s]
        for k in ("itU", "itV", "itP", "iter"):
            assert info[k] == sg[k], (step, k, info[k], sg[k])
        for k in ("error", "residU", "residV", "residMass"):
            assert info[k] == sg[k], (step, k, info[k], sg[k])
        its.append(info["itP"])
    if ptol is not None:
        # the CG stopped inside its budget in some step (the case this covers)
        assert any(0 < x < pcg for x in its), its
    u, v, pp = b.fields(s)
    np.testing.assert_array_equal(u, g.field(cb.F_U))
    np.testing.assert_array_equal(v, g.field(cb.F_V))
    np.testing.assert_array_equal(pp, g.field(cb.F_P))
    b.close()
    g.close()


@pytest.mark.parametrize("precision", [cb.CP_PREC_FP64, cb.CP_PREC_MIXED], ids=["fp64", "mixed"])
def test_failed_slot_is_isolated(precision):
    """A slot whose state turns non-finite ends CP_SLOT_FAILED with
    CP_E_DIVERGED at the end of that outer iteration. The slots packed with
    it stay bit-identical to the same geometries run alone (every reduction
    and every skip rule is per slot), and its arena space is reused."""
    gs = []
    for p, ct in geoms():
        q = wl.fast_mode(p) if precision == cb.CP_PREC_MIXED else p.copy()
        q.precision = precision
        gs.append((q, ct))
    say(f"[batch] failure isolation, precision {precision}")
    alone = {}
    for k in (0, 2):
        b = cb.Batch(precision, 256 << 20, 1)
        s = b.admit(*gs[k])
        b.run(20)
        alone[k] = b.fields(s)
        b.close()
    b = cb.Batch(precision, 256 << 20, 4)
    slots = [b.admit(*g) for g in gs]
    b.run(8)
    u, _, _ = b.fields(slots[1])
    u[u.shape[0] // 2, u.shape[1] // 2] = np.nan
    b.set_fields(slots[1], u=u)
    b.run(12)
    info = b.poll()
    bad = info[slots[1]]
    assert bad["status"] == cb.SLOT_FAILED and bad["health"] == cb.CP_E_DIVERGED, bad
    assert bad["iter"] == 9, bad  # the iteration that read the NaN ends it
    for k in (0, 2):
        assert info[slots[k]]["iter"] == 20 and info[slots[k]]["status"] == cb.SLOT_RUNNING
        for x, y in zip(b.fields(slots[k]), alone[k]):
            np.testing.assert_array_equal(x, y)
    b.evict(slots[1])
    s_new = b.admit(*gs[1])
    assert s_new == slots[1]
    b.run(3)
    info = b.poll()
    assert info[s_new]["status"] == cb.SLOT_RUNNING and info[s_new]["iter"] == 3 and info[s_new]["health"] == 0
    assert info[slots[0]]["iter"] == 23
    b.close()


@pytest.mark.parametrize("precision", [cb.CP_PREC_FP64, cb.CP_PREC_MIXED], ids=["fp64", "mixed"])
def test_graph_batch_mixed_budgets_equal_alone(precision):
    """One graph batch holding slots whose Jacobi, CG and refinement budgets
    differ: the captured step takes the largest budget of each loop, and
    every slot stops its own loops at its own budget. Step by step, each
    slot's inner counts and error, and its final fields bit for bit, equal
    the same slot run alone in a batch of its own."""
    base = [wl.network(64, 128, 3, precision), wl.network(48, 128, 11, precision),
            wl.gamma_field(40, 96, 5, precision, corr=5.0)]
    # (Jacobi sweeps, CG iterations, refinement cycles)
    budgets = [(2, 3, 1), (4, 5, 2), (3, 2, 1)]
    gs = []
    for (p, ct), (js, ps, rc) in zip(base, budgets):
        q = wl.fast_mode(p, js, ps, rc) if precision == cb.CP_PREC_MIXED else p.copy()
        q.precision = precision
        q.pressurePrecond = cb.CP_PRECOND_AMG
        if precision == cb.CP_PREC_FP64:
            q.maxMomentumIter, q.maxPressureIter = js, ps
        gs.append((q, ct))
    say(f"[batch] mixed budgets in one graph, precision {precision}")
    steps = 8
    ref = []
    for q, ct in gs:
        b = cb.Batch(precision, 128 << 20, 1)
        s = b.admit(q, ct)
        hist = []
        for _ in range(steps):
            b.run(1)
            i = b.poll()[s]
            hist.append((i["itU"], i["itV"], i["itP"], i["error"]))
        ref.append((hist, b.fields(s)))
        b.close()
    b = cb.Batch(precision, 256 << 20, 3)
    slots = [b.admit(q, ct) for q, ct in gs]
    for step in range(steps):
        b.run(1)
        info = b.poll()
        for k, s in enumerate(slots):
            got = (info[s]["itU"], info[s]["itV"], info[s]["itP"], info[s]["error"])
            assert got == ref[k][0][step], (k, step, got, ref[k][0][step])
    for k, s in enumerate(slots):
        for x, y in zip(b.fields(s), ref[k][1]):
            np.testing.assert_array_equal(x, y)
    b.close()
