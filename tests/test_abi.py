"""C ABI checks that need no GPU: the library loads, exports every symbol
declared in include/*.h, parameter defaults match the reference, host-only
entry points behave, and errors come back as codes (no GPU calls)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

import cpbind as cb
import workloads as wl

REPO = Path(__file__).resolve().parent.parent


def declared_symbols() -> set[str]:
    names = set()
    for h in (REPO / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(cp_\w+)\s*\(", text, flags=re.M):
            if not m.group(1).startswith("cp_fill_"):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = cb.load()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in cb.EXPORTS}
    assert names == bound, (names ^ bound)


def test_params_struct_layout_matches_header():
    text = (REPO / "include" / "coldplate_hip.h").read_text()
    body = text[text.index("typedef struct cp_params {") + len("typedef struct cp_params {"):text.index("} cp_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl or decl.startswith("typedef"):
            continue
        parts = decl.split(None, 1)
        fields += [f.strip() for f in parts[1].split(",")]
    assert [f for f, _ in cb.Params._fields_] == fields


def test_reference_defaults():
    p = cb.default_params()
    # SIMPLE.h:68-358
    assert p.epsilon == 1e-5 and p.maxIterations == 10000
    assert p.maxPressureIter == 20 and p.pTol == 1e-4
    assert p.momentumTol == 1e-5 and p.maxMomentumIter == 200
    assert p.uvAlpha == 0.7 and p.pAlpha == 0.2
    assert p.rho == 997.0 and p.mu == 0.00089
    assert p.sinkMultiplier == 4.8 and p.KMin == 1e-14 and p.brinkmanQ == 1.0
    assert p.enableSER == 1 and p.serCFLMax == 5.0 and p.serMinIter == 10
    assert p.pseudoCFLInitial == 0.1 and p.pseudoCFLMax == 30.0
    assert p.precision == cb.CP_PREC_FP64 and p.massResidualScope == 0


def test_create_rejects_bad_arguments_without_gpu_work():
    lib = cb.load()
    p = cb.default_params()
    p.M, p.N = 2, 2
    h = C.c_void_p()
    ct = np.zeros((2, 2))
    rc = lib.cp_create(C.byref(p), cb.dptr(ct), C.byref(h))
    assert rc == -1
    assert b"at least 4" in lib.cp_last_error(None)


def test_thermal_defaults_and_argument_errors_need_no_gpu():
    p = cb.thermal_default_params()
    # struct Params, ThermalSolver/thermal_solver.cpp:142-189
    assert (p.nz_solid, p.nz_fluid, p.H_b, p.H_t) == (2, 7, 0.0005, 0.004)
    assert (p.k_s, p.k_f, p.rho, p.Cp, p.T_inlet) == (400.0, 0.6, 998.0, 4180.0, 30.0)
    assert (p.rtol, p.maxiter, p.k_interp_mode, p.qk) == (1e-6, 1000, 1, 1.0)
    lib = cb.load()
    g = np.ones((4, 4))
    T = np.empty((9, 4, 4))
    st = cb.ThermalStats()
    for field, val, msg in (("Nx", 0, b"grid"), ("nz_fluid", 20, b"layers"), ("dx", 0.0, b"dx"),
                            ("k_interp_mode", 3, b"k_interp_mode"), ("rtol", 0.0, b"rtol")):
        q = cb.thermal_default_params()
        q.Nx, q.Ny, q.dx, q.dy, q.q0 = 4, 4, 1e-4, 1e-4, 1e5
        setattr(q, field, val)
        rc = lib.cp_thermal_solve(C.byref(q), cb.dptr(g), cb.dptr(g), cb.dptr(g), cb.dptr(T), C.byref(st))
        assert rc == -1, field
        assert msg in lib.cp_last_error(None), field


def test_thermal_structs_match_header():
    text = re.sub(r"/\*.*?\*/", "", (REPO / "include" / "coldplate_hip.h").read_text(), flags=re.S)
    for name, cls in (("cp_thermal_params", cb.ThermalParams), ("cp_thermal_stats", cb.ThermalStats)):
        body = text[text.index(f"typedef struct {name} {{") + len(f"typedef struct {name} {{"):text.index(f"}} {name};")]
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if decl:
                fields += [f.strip() for f in decl.split(None, 1)[1].split(",")]
        assert [f for f, _ in cls._fields_] == fields, name


def test_geometry_generators_deterministic_and_valid():
    a = wl.geom_network(128, 256, 5)
    b = wl.geom_network(128, 256, 5)
    c = wl.geom_network(128, 256, 6)
    np.testing.assert_array_equal(a, b)
    assert np.any(a != c)
    assert set(np.unique(a)) <= {0.0, 1.0}
    assert np.all(a[0] == 1) and np.all(a[-1] == 1)
    assert (a[:, 0] == 0).sum() >= 8 and (a[:, -1] == 0).sum() >= 8
    g = wl.geom_gamma(96, 128, 3, corr=8.0)
    np.testing.assert_array_equal(g, wl.geom_gamma(96, 128, 3, corr=8.0))
    assert g.min() >= 0 and g.max() <= 1
    assert 0.2 < (g < 0.5).mean() < 0.8  # roughly half fluid
    s = wl.geom_straight(50, 200, 10, 39)
    assert s[10:40].sum() == 0 and s[:10].min() == 1 and s[40:].min() == 1


def test_stream_and_batch_workload_shapes():
    reqs = wl.stream_2000(50)
    assert len(reqs) == 50
    for r in reqs:
        assert 32 <= r["M"] <= 1024 and 32 <= r["N"] <= 2048
    b64 = wl.batch_64()
    assert len(b64) == 64
    for p, ct in b64:
        assert ct.shape == (p.M, p.N)
        assert 128 <= p.N <= 1024 and p.M == max(16, (p.N // 2 // 8) * 8)


def test_serve_pool_argument_errors_need_no_gpu():
    lib = cb.load()
    assert lib.cp_serve_pool(None, 0, None, 0, 1, None) == cb.CP_E_ARG
    hs = (C.c_void_p * 2)(None, None)
    assert lib.cp_serve_pool(C.cast(hs, C.c_void_p), 2, None, 0, 1, None) == cb.CP_E_ARG
