"""ctypes binding of the C ABI declared in include/coldplate_hip.h.

The structures below mirror the C structs field for field; the library is
``coldplate-solver_amd/lib/libcoldplate_hip.so`` built by ``build.py``.
Loading fails loudly when the library is missing: there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libcoldplate_hip.so"

CP_OK = 0
CP_E_ARG, CP_E_HIP, CP_E_NOMEM, CP_E_STATE, CP_E_CAPACITY, CP_E_NODEV, CP_E_DIVERGED = range(-1, -8, -1)
CP_PREC_FP64 = 0
CP_PREC_MIXED = 1
CP_PRECOND_JACOBI = 0
CP_PRECOND_AMG = 1

F_U, F_V, F_P, F_USTAR, F_VSTAR, F_PSTAR, F_DE, F_DN, F_GAMMA, F_ALPHA = range(10)

SLOT_FREE, SLOT_RUNNING, SLOT_CONVERGED, SLOT_MAXITER, SLOT_FAILED = range(5)

K_MOM_ASM = 0
K_PCG_Q = 4
K_MG_UP = 15


class Params(C.Structure):
    """cp_params (include/coldplate_hip.h)."""

    _fields_ = [
        ("M", C.c_int32), ("N", C.c_int32),
        ("hx", C.c_double), ("hy", C.c_double),
        ("rho", C.c_double), ("mu", C.c_double),
        ("U_in", C.c_double),
        ("enableInletRamp", C.c_int32), ("rampSteps", C.c_int32),
        ("uvAlpha", C.c_double), ("pAlpha", C.c_double),
        ("convectionScheme", C.c_int32),
        ("enableTwoPointFiveD", C.c_int32), ("enableConvectionScaling", C.c_int32),
        ("Ht", C.c_double), ("sinkMultiplier", C.c_double),
        ("alphaMax", C.c_double), ("KMin", C.c_double), ("brinkmanQ", C.c_double),
        ("enablePseudoTime", C.c_int32), ("useLocalPseudoTime", C.c_int32),
        ("timeStepMultiplier", C.c_double), ("minPseudoSpeed", C.c_double),
        ("enableCflRamp", C.c_int32),
        ("pseudoCFLInitial", C.c_double), ("pseudoCFLMax", C.c_double),
        ("cflRampStartRes", C.c_double), ("cflRampExponent", C.c_double),
        ("cflRampSmooth", C.c_double),
        ("enableSER", C.c_int32), ("serMinIter", C.c_int32), ("serUseMaxResid", C.c_int32),
        ("serCFLMin", C.c_double), ("serCFLMax", C.c_double), ("serSmooth", C.c_double),
        ("serMaxIncrease", C.c_double), ("serMaxDecrease", C.c_double),
        ("enableLineSearch", C.c_int32), ("lsMaxTries", C.c_int32),
        ("lsAlphaMin", C.c_double), ("lsAlphaReduce", C.c_double),
        ("lsResidIncreaseTol", C.c_double),
        ("momentumTol", C.c_double), ("maxMomentumIter", C.c_int32),
        ("pTol", C.c_double), ("maxPressureIter", C.c_int32),
        ("epsilon", C.c_double), ("maxIterations", C.c_int32),
        ("precision", C.c_int32),
        ("refineCycles", C.c_int32),
        ("fixedMomentumSweeps", C.c_int32),
        ("massResidualScope", C.c_int32),
        ("pressurePrecond", C.c_int32),
        ("mgOmega", C.c_double),
        ("mgScale", C.c_double),
        ("usePressureDropConvergence", C.c_int32), ("dpConvergenceWindow", C.c_int32),
        ("dpConvergencePercent", C.c_double),
        ("usePressureDropSlopeGate", C.c_int32), ("dpSlopeWindowIters", C.c_int32),
        ("dpSlopeMaxDelta", C.c_double),
        ("xPlaneCoreInlet", C.c_double), ("xPlaneCoreOutlet", C.c_double),
        ("pressureSolverType", C.c_int32), ("momentumSolverType", C.c_int32),
        ("sorOmega", C.c_double),
    ]

    def copy(self) -> "Params":
        q = Params()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(Params))
        return q

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepStats(C.Structure):
    """cp_step_stats."""

    _fields_ = [
        ("error", C.c_double), ("residU", C.c_double), ("residV", C.c_double),
        ("residMass", C.c_double), ("pseudoCFL", C.c_double),
        ("itU", C.c_int32), ("itV", C.c_int32), ("itP", C.c_int32),
        ("iter", C.c_int32), ("converged", C.c_int32), ("health", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class PlaneMetrics(C.Structure):
    """cp_plane_metrics (PlaneMetrics, SIMPLE.h:679-686)."""

    _fields_ = [
        ("flowArea", C.c_double), ("massFlux", C.c_double), ("avgStatic", C.c_double),
        ("avgDynamic", C.c_double), ("avgTotal", C.c_double), ("valid", C.c_int32),
    ]


class ThermalParams(C.Structure):
    """cp_thermal_params (include/coldplate_hip.h)."""

    _fields_ = [
        ("Nx", C.c_int32), ("Ny", C.c_int32), ("nz_solid", C.c_int32), ("nz_fluid", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("H_b", C.c_double), ("H_t", C.c_double),
        ("k_s", C.c_double), ("k_f", C.c_double), ("rho", C.c_double), ("Cp", C.c_double),
        ("T_inlet", C.c_double), ("q0", C.c_double), ("k_interp_mode", C.c_int32), ("qk", C.c_double),
        ("rtol", C.c_double), ("maxiter", C.c_int32), ("timing", C.c_int32),
    ]

    def copy(self) -> "ThermalParams":
        q = ThermalParams()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(ThermalParams))
        return q


class ThermalStats(C.Structure):
    """cp_thermal_stats."""

    _fields_ = [
        ("iters", C.c_int32), ("converged", C.c_int32), ("relres", C.c_double),
        ("T_base_min", C.c_double), ("T_base_max", C.c_double), ("T_fin_min", C.c_double),
        ("T_fin_max", C.c_double), ("T_fluid_min", C.c_double), ("T_fluid_max", C.c_double),
        ("T_avg_base", C.c_double), ("T_avg_fluid", C.c_double), ("T_avg_solid", C.c_double),
        ("T_avg_global", C.c_double), ("Q_in", C.c_double), ("Q_out", C.c_double), ("T_out", C.c_double),
        ("imbalance_pct", C.c_double), ("device_ms", C.c_double), ("sweep_ms", C.c_double),
        ("sweeps", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Request(C.Structure):
    _fields_ = [("prm", C.POINTER(Params)), ("cellType", C.POINTER(C.c_double)), ("arrival", C.c_double),
                ("gammaMin", C.c_double)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("iter", C.c_int32), ("slot", C.c_int32), ("rc", C.c_int32),
                ("error", C.c_double), ("dp", C.c_double), ("tAdmit", C.c_double), ("tDone", C.c_double)]


class SlotInfo(C.Structure):
    """cp_slot_info."""

    _fields_ = [
        ("status", C.c_int32), ("iter", C.c_int32),
        ("error", C.c_double), ("residU", C.c_double), ("residV", C.c_double),
        ("residMass", C.c_double), ("health", C.c_int32), ("M", C.c_int32), ("N", C.c_int32),
        ("itU", C.c_int32), ("itV", C.c_int32), ("itP", C.c_int32), ("pseudoCFL", C.c_double),
    ]


def field_shape(which: int, M: int, N: int) -> tuple[int, int]:
    """Reference shape of a CP_F_* field (SIMPLE.cpp:184-201)."""
    if which in (F_U, F_USTAR, F_DE):
        return (M + 1, N)
    if which in (F_V, F_VSTAR, F_DN):
        return (M, N + 1)
    if which in (F_P, F_PSTAR):
        return (M + 1, N + 1)
    return (M, N)


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


# (name, restype, argtypes) of every exported symbol of include/coldplate_hip.h.
EXPORTS = [
    ("cp_params_default", C.c_int, [C.POINTER(Params)]),
    ("cp_last_error", C.c_char_p, [_vp]),
    ("cp_device_count", C.c_int, []),
    ("cp_create", C.c_int, [C.POINTER(Params), _dp, C.POINTER(_vp)]),
    ("cp_destroy", C.c_int, [_vp]),
    ("cp_set_fields", C.c_int, [_vp, _dp, _dp, _dp]),
    ("cp_get_fields", C.c_int, [_vp, _dp, _dp, _dp]),
    ("cp_get_field", C.c_int, [_vp, C.c_int, _dp]),
    ("cp_momentum", C.c_int, [_vp, C.POINTER(StepStats)]),
    ("cp_pressure", C.c_int, [_vp, C.POINTER(StepStats)]),
    ("cp_correct", C.c_int, [_vp]),
    ("cp_step", C.c_int, [_vp, C.POINTER(StepStats)]),
    ("cp_run", C.c_int, [_vp, C.c_int, C.POINTER(StepStats)]),
    ("cp_plane", C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(PlaneMetrics)]),
    ("cp_sync", C.c_int, [_vp]),
    ("cp_batch_create", C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(_vp)]),
    ("cp_batch_destroy", C.c_int, [_vp]),
    ("cp_batch_footprint", C.c_size_t, [_vp, C.c_int, C.c_int]),
    ("cp_batch_admit", C.c_int, [_vp, C.POINTER(Params), _dp, C.POINTER(C.c_int)]),
    ("cp_batch_run", C.c_int, [_vp, C.c_int]),
    ("cp_batch_poll", C.c_int, [_vp, C.POINTER(SlotInfo)]),
    ("cp_batch_get_fields", C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    ("cp_batch_set_fields", C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    ("cp_batch_plane", C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.POINTER(PlaneMetrics)]),
    ("cp_batch_evict", C.c_int, [_vp, C.c_int]),
    ("cp_batch_set_tiling", C.c_int, [_vp, C.c_int, C.c_int]),
    ("cp_batch_set_graph", C.c_int, [_vp, C.c_int]),
    ("cp_batch_set_mg_tiling", C.c_int, [_vp, C.c_int]),
    ("cp_serve", C.c_int, [_vp, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    ("cp_serve_pool", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    ("cp_batch_autotune", C.c_int, [_vp, C.c_int, _ip, _ip, _ip, _ip, _dp]),
    ("cp_batch_set_tile_order", C.c_int, [_vp, C.c_int]),
    ("cp_run_noalloc", C.c_int, [C.POINTER(Params), _dp, _dp, _dp, _dp, C.POINTER(StepStats)]),
    ("cp_thermal_params_default", C.c_int, [C.POINTER(ThermalParams)]),
    ("cp_thermal_solve", C.c_int, [C.POINTER(ThermalParams), _dp, _dp, _dp, _dp, C.POINTER(ThermalStats)]),
    ("cp_io_read_params", C.c_int, [C.c_char_p, C.POINTER(Params), _ip, _ip]),
    ("cp_io_read_geometry", C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp]),
    ("cp_io_read_matrix", C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp]),
    ("cp_io_write_matrix", C.c_int, [C.c_char_p, _dp, C.c_int, C.c_int, C.c_int]),
    ("cp_io_save_all", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                 _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]),
    ("cp_batch_probe", C.c_int, [_vp, C.c_int]),
    ("cp_batch_probe_read", C.c_int, [_vp, _ip, _ip, _dp, _dp, C.c_int]),
    ("cp_batch_probe_burst", C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    ("cp_geom_straight", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    ("cp_geom_network", C.c_int, [C.c_int, C.c_int, C.c_uint64, _dp]),
    ("cp_geom_gamma", C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double, _dp]),
]

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python __graft_entry__.py build)")
    lib = C.CDLL(str(p))
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class CpError(RuntimeError):
    pass


def check(rc: int, ctx=None) -> None:
    if rc != CP_OK:
        msg = load().cp_last_error(ctx)
        raise CpError(f"cp call failed ({rc}): {msg.decode() if msg else ''}")


def default_params() -> Params:
    p = Params()
    load().cp_params_default(C.byref(p))
    return p


class Solver:
    """Single-geometry device solve (cp_ctx)."""

    def __init__(self, prm: Params, cell_type: np.ndarray):
        lib = load()
        ct = np.ascontiguousarray(cell_type, dtype=np.float64)
        assert ct.shape == (prm.M, prm.N)
        self.prm = prm.copy()
        self.M, self.N = prm.M, prm.N
        h = _vp()
        check(lib.cp_create(C.byref(self.prm), dptr(ct), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().cp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stats(self, fn) -> dict:
        st = StepStats()
        check(fn(self.h, C.byref(st)), self.h)
        return st.as_dict()

    def momentum(self) -> dict:
        return self._stats(load().cp_momentum)

    def pressure(self) -> dict:
        return self._stats(load().cp_pressure)

    def correct(self) -> None:
        check(load().cp_correct(self.h), self.h)

    def step(self) -> dict:
        return self._stats(load().cp_step)

    def run(self, check_every: int = 0) -> dict:
        st = StepStats()
        check(load().cp_run(self.h, check_every, C.byref(st)), self.h)
        return st.as_dict()

    def sync(self) -> None:
        check(load().cp_sync(self.h), self.h)

    def field(self, which: int) -> np.ndarray:
        out = np.empty(field_shape(which, self.M, self.N), dtype=np.float64)
        check(load().cp_get_field(self.h, which, dptr(out)), self.h)
        return out

    def set_fields(self, u=None, v=None, p=None) -> None:
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, p)]
        check(load().cp_set_fields(self.h, *[dptr(a) for a in arrs]), self.h)

    def plane(self, x: float, gamma_min: float = 1.0 - 1e-6) -> dict:
        m = PlaneMetrics()
        check(load().cp_plane(self.h, x, gamma_min, C.byref(m)), self.h)
        return {k: getattr(m, k) for k, _ in m._fields_}

    def save_all(self, directory, cell_type: np.ndarray, n_in: int = 0, n_out: int = 0, precision: int = 0) -> None:
        """saveAll's files (u/v/p, cell-centred, thermal crops, VTK)."""
        save_all(directory, self.prm, self.field(F_U), self.field(F_V), self.field(F_P), cell_type,
                 self.field(F_GAMMA), self.field(F_ALPHA), n_in, n_out, precision)

    @classmethod
    def from_files(cls, params_path, geometry_path, base: "Params | None" = None) -> "tuple[Solver, np.ndarray]":
        """A solve from the reference's fluid_params.txt / geometry_fluid.txt
        (other parameters from `base`, else the reference defaults)."""
        prm, _, _ = read_params(params_path, base)
        ct, _ = read_geometry(geometry_path, prm.M, prm.N)
        return cls(prm, ct), ct

    def full_dp(self, gamma_min: float = 1.0 - 1e-6) -> float:
        """Full-system total-pressure drop (SIMPLE.cpp:343-349)."""
        a = self.plane(0.0, gamma_min)
        b = self.plane(self.N * self.prm.hx, gamma_min)
        return a["avgTotal"] - b["avgTotal"]


def _request_array(requests):
    """cp_request array of [(params, cellType, arrival_s, gamma_min), ...],
    with the Python objects it points into (keep them alive for the call)."""
    n = len(requests)
    keep = []
    reqs = (Request * max(n, 1))()
    for k, (prm, ct, arrival, gmin) in enumerate(requests):
        pr = prm.copy()
        ca = np.ascontiguousarray(ct, dtype=np.float64)
        keep.append((pr, ca))
        reqs[k].prm = C.pointer(pr)
        reqs[k].cellType = ca.ctypes.data_as(C.POINTER(C.c_double))
        reqs[k].arrival = float(arrival)
        reqs[k].gammaMin = float(gmin)
    return reqs, keep, n


def serve_pool(batches, requests, steps_per_poll: int = 20) -> list[dict]:
    """cp_serve_pool: `requests` served by several empty batches at once, one
    host thread and HIP stream each, from one first-come-first-served queue;
    one result dict per request, in order."""
    reqs, keep, n = _request_array(requests)
    hs = (C.c_void_p * max(len(batches), 1))(*[b.h for b in batches])
    res = (Result * max(n, 1))()
    check(load().cp_serve_pool(C.cast(hs, C.c_void_p), len(batches), C.cast(reqs, C.c_void_p), n, steps_per_poll,
                               C.cast(res, C.c_void_p)))
    del keep
    return [{f: getattr(res[k], f) for f, _ in Result._fields_} for k in range(n)]


class Batch:
    """Packed multi-geometry device solve (cp_batch)."""

    def __init__(self, precision: int, hbm_budget: int, max_slots: int):
        lib = load()
        h = _vp()
        check(lib.cp_batch_create(precision, hbm_budget, max_slots, C.byref(h)))
        self.h = h
        self.max_slots = max_slots
        self.shapes: dict[int, tuple[int, int, float]] = {}

    def close(self):
        if getattr(self, "h", None):
            load().cp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def footprint(self, M: int, N: int) -> int:
        return int(load().cp_batch_footprint(self.h, M, N))

    def admit(self, prm: Params, cell_type: np.ndarray) -> int | None:
        """Returns the slot, or None when the batch is full."""
        ct = np.ascontiguousarray(cell_type, dtype=np.float64)
        s = C.c_int(-1)
        rc = load().cp_batch_admit(self.h, C.byref(prm), dptr(ct), C.byref(s))
        if rc == -5:
            return None
        check(rc, self.h)
        self.shapes[s.value] = (prm.M, prm.N, prm.hx)
        return s.value

    def run(self, steps: int) -> None:
        check(load().cp_batch_run(self.h, steps), self.h)

    def poll(self) -> list[dict]:
        info = (SlotInfo * self.max_slots)()
        check(load().cp_batch_poll(self.h, info), self.h)
        return [{k: getattr(x, k) for k, _ in x._fields_} for x in info]

    def fields(self, slot: int):
        M, N, _ = self.shapes[slot]
        u = np.empty((M + 1, N)); v = np.empty((M, N + 1)); p = np.empty((M + 1, N + 1))
        check(load().cp_batch_get_fields(self.h, slot, dptr(u), dptr(v), dptr(p)), self.h)
        return u, v, p

    def set_fields(self, slot: int, u=None, v=None, p=None) -> None:
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, p)]
        check(load().cp_batch_set_fields(self.h, slot, *[dptr(a) for a in arrs]), self.h)

    def plane(self, slot: int, x: float, gamma_min: float = 1.0 - 1e-6) -> dict:
        m = PlaneMetrics()
        check(load().cp_batch_plane(self.h, slot, x, gamma_min, C.byref(m)), self.h)
        return {k: getattr(m, k) for k, _ in m._fields_}

    def full_dp(self, slot: int, gamma_min: float = 1.0 - 1e-6) -> float:
        M, N, hx = self.shapes[slot]
        return self.plane(slot, 0.0, gamma_min)["avgTotal"] - self.plane(slot, N * hx, gamma_min)["avgTotal"]

    def set_tiling(self, tile_rows: int, wg_per_cu: int = 0) -> None:
        check(load().cp_batch_set_tiling(self.h, tile_rows, wg_per_cu), self.h)

    def set_graph(self, use_graph: bool) -> None:
        check(load().cp_batch_set_graph(self.h, 1 if use_graph else 0), self.h)

    def serve(self, requests, steps_per_poll: int = 20) -> list[dict]:
        """Continuous batching over `requests` = [(params, cellType, arrival_s,
        gamma_min), ...]; returns one result dict per request, in order."""
        reqs, keep, n = _request_array(requests)
        res = (Result * max(n, 1))()
        check(load().cp_serve(self.h, C.cast(reqs, C.c_void_p), n, steps_per_poll, C.cast(res, C.c_void_p)), self.h)
        del keep
        return [{f: getattr(res[k], f) for f, _ in Result._fields_} for k in range(n)]

    def autotune(self, steps: int = 2) -> dict:
        """Time every candidate tiling, then every tile order, on the admitted
        slots; keep the fastest."""
        r, m, w, o = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
        ms = C.c_double(0.0)
        check(load().cp_batch_autotune(self.h, steps, C.byref(r), C.byref(m), C.byref(w), C.byref(o), C.byref(ms)),
              self.h)
        return {"tile_rows": r.value, "mg_rows": m.value, "wg_per_cu": w.value, "tile_order": o.value,
                "ms_per_step": ms.value}

    def set_tile_order(self, order: int) -> None:
        check(load().cp_batch_set_tile_order(self.h, order), self.h)

    def set_mg_tiling(self, rows: int) -> None:
        check(load().cp_batch_set_mg_tiling(self.h, rows), self.h)

    def probe(self, kernels=()) -> None:
        """Time the kernels (K_PCG_Q, K_MOM_ASM, K_MG_UP) in bursts during serve."""
        mask = 0
        for k in kernels:
            mask |= 1 << k
        check(load().cp_batch_probe(self.h, mask), self.h)

    def probe_read(self, cap: int = 65536) -> list[dict]:
        """Timing bursts of the last serve: kernel, launches, elapsed ms and
        work units over the burst's launches."""
        kern = np.zeros(cap, np.int32)
        nl = np.zeros(cap, np.int32)
        ms = np.zeros(cap)
        un = np.zeros(cap)
        n = load().cp_batch_probe_read(self.h, kern.ctypes.data_as(_ip), nl.ctypes.data_as(_ip), dptr(ms), dptr(un),
                                       cap)
        if n < 0:
            check(n, self.h)
        return [dict(kernel=int(kern[k]), launches=int(nl[k]), ms=float(ms[k]), units=float(un[k])) for k in range(n)]

    def probe_burst(self, launches: int, kernel: int = K_PCG_Q) -> float:
        """Mean ms per launch of `launches` back-to-back k_pcg_q launches."""
        ms = C.c_double(0.0)
        check(load().cp_batch_probe_burst(self.h, kernel, launches, C.byref(ms)), self.h)
        return ms.value

    def evict(self, slot: int) -> None:
        check(load().cp_batch_evict(self.h, slot), self.h)
        self.shapes.pop(slot, None)


# ---------------------------------------------------------------------------
# the reference's file formats
# ---------------------------------------------------------------------------
def read_params(path, base: Params | None = None) -> tuple[Params, int, int]:
    """fluid_params.txt -> (params, N_in_buffer, N_out_buffer)."""
    p = base.copy() if base is not None else default_params()
    ni, no = C.c_int32(0), C.c_int32(0)
    check(load().cp_io_read_params(str(path).encode(), C.byref(p), C.byref(ni), C.byref(no)))
    return p, ni.value, no.value


def read_geometry(path, M: int, N: int) -> tuple[np.ndarray, bool]:
    """geometry_fluid.txt -> (cellType M x N, file_found)."""
    ct = np.zeros((M, N))
    rc = load().cp_io_read_geometry(str(path).encode(), M, N, dptr(ct))
    if rc < 0:
        check(rc)
    return ct, rc == 0


def read_matrix(path, rows: int, cols: int) -> np.ndarray:
    a = np.zeros((rows, cols))
    check(load().cp_io_read_matrix(str(path).encode(), rows, cols, dptr(a)))
    return a


def write_matrix(path, a: np.ndarray, precision: int = 0) -> None:
    a = np.ascontiguousarray(a, dtype=np.float64)
    check(load().cp_io_write_matrix(str(path).encode(), dptr(a), a.shape[0], a.shape[1], precision))


def save_all(directory, prm: Params, u, v, p, cell_type, gamma, alpha, n_in: int = 0, n_out: int = 0,
             precision: int = 0) -> None:
    """saveAll's files for fields in the reference shapes."""
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (u, v, p, cell_type, gamma, alpha)]
    check(load().cp_io_save_all(str(directory).encode(), prm.M, prm.N, prm.hx, prm.hy, n_in, n_out,
                                 *(dptr(x) for x in arrs), precision))


# ---------------------------------------------------------------------------
# downstream conjugate heat transfer
# ---------------------------------------------------------------------------
def thermal_default_params() -> ThermalParams:
    p = ThermalParams()
    check(load().cp_thermal_params_default(C.byref(p)))
    return p


def thermal_inputs(u: np.ndarray, v: np.ndarray, cell_type: np.ndarray, n_in: int = 0, n_out: int = 0):
    """The thermal solve's inputs from a SIMPLE result, as the reference's
    files hold them: u_thermal / v_thermal are the cell-centred velocities
    (Utilities/output.cpp:87-91) of columns n_in .. N - n_out - 1
    (output.cpp:155-181), and gamma = 1 - cellType over the same columns
    (geometry_thermal.txt, inverted by thermal_solver.cpp:738-743)."""
    M, N = cell_type.shape
    assert u.shape == (M + 1, N) and v.shape == (M, N + 1)
    assert 0 <= n_in and 0 <= n_out and n_in + n_out < N
    uc = 0.5 * (u[:-1, :] + u[1:, :])
    vc = 0.5 * (v[:, :-1] + v[:, 1:])
    cols = slice(n_in, N - n_out)
    return (np.ascontiguousarray(1.0 - cell_type[:, cols]), np.ascontiguousarray(uc[:, cols]),
            np.ascontiguousarray(vc[:, cols]))


def thermal_solve(prm: ThermalParams, gamma: np.ndarray, u: np.ndarray, v: np.ndarray):
    """cp_thermal_solve: (T of shape (Nz, Ny, Nx), statistics dict)."""
    g, uu, vv = (np.ascontiguousarray(a, dtype=np.float64) for a in (gamma, u, v))
    assert g.shape == uu.shape == vv.shape == (prm.Ny, prm.Nx)
    T = np.empty((prm.nz_solid + prm.nz_fluid, prm.Ny, prm.Nx))
    st = ThermalStats()
    pr = prm.copy()
    rc = load().cp_thermal_solve(C.byref(pr), dptr(g), dptr(uu), dptr(vv), dptr(T), C.byref(st))
    check(rc)
    return T, st.as_dict()
