"""Workload definitions: solver parameters + seeded geometries.

Each function returns ``(Params, cellType)``. Sizes follow the five
workloads the backend is measured on:

* straight_200x50      200 x 50 straight channel (rows 10-39 fluid), FOU,
                       FP64, the CPU-runnable case.
* network_1024x512     seeded channel network, 2.5D model (Ht = 200 um),
                       SOU, pseudo-CFL 0.5 -> 5, plain FP64.
* gamma_2048           2048 x 2048 tanh-projected Gaussian density field,
                       alpha_max = 1e9, 2.5D sink, mixed precision.
* batch_64             64 geometries, 256x128 ... 1024x512 log-uniform,
                       half networks, half gamma fields, mixed precision.
* stream_2000          2000 requests, 128^2 ... 2048x1024, straight /
                       network / gamma, 2.5D on/off.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from cpbind import CP_PREC_FP64, CP_PREC_MIXED, Params, dptr, load

WATER_RHO = 1000.0
WATER_MU = 1e-3
GAMMA_PLANE_MIN = 0.5  # plane-row threshold for continuous gamma fields


def base_params(M: int, N: int, h: float, U: float) -> Params:
    p = Params()
    load().cp_params_default(C.byref(p))
    p.M, p.N = M, N
    p.hx = p.hy = h
    p.rho, p.mu = WATER_RHO, WATER_MU
    p.U_in = U
    p.uvAlpha, p.pAlpha = 0.7, 0.3
    p.epsilon = 1e-6
    p.maxIterations = 200000
    # the configurations converge on the residual stop test alone
    # (usePressureDropConvergence = false)
    p.usePressureDropConvergence = 0
    return p


def geom_straight(M: int, N: int, r0: int, r1: int) -> np.ndarray:
    ct = np.empty((M, N))
    rc = load().cp_geom_straight(M, N, r0, r1, dptr(ct))
    assert rc == 0
    return ct


def geom_network(M: int, N: int, seed: int) -> np.ndarray:
    ct = np.empty((M, N))
    rc = load().cp_geom_network(M, N, seed, dptr(ct))
    assert rc == 0
    return ct


def geom_gamma(M: int, N: int, seed: int, corr: float = 32.0, beta: float = 8.0, eta: float = 0.5) -> np.ndarray:
    ct = np.empty((M, N))
    rc = load().cp_geom_gamma(M, N, seed, corr, beta, eta, dptr(ct))
    assert rc == 0
    return ct


def straight_200x50(precision: int = CP_PREC_FP64):
    """Plain method, 200 x 50, 10 mm x 2.5 mm, channel on cell rows 10-39."""
    p = base_params(50, 200, 5e-5, 0.1)
    p.precision = precision
    return p, geom_straight(50, 200, 10, 39)


def network_params(M: int, N: int, U: float = 0.2, h: float = 25e-6) -> Params:
    p = base_params(M, N, h, U)
    p.enableTwoPointFiveD = 1
    p.enableConvectionScaling = 1
    p.Ht = 2e-4
    p.convectionScheme = 1
    p.enablePseudoTime = 1
    p.pseudoCFLInitial = 0.5
    p.serCFLMax = 5.0
    return p


def network(M: int, N: int, seed: int, precision: int = CP_PREC_FP64, U: float = 0.2):
    p = network_params(M, N, U)
    p.precision = precision
    return p, geom_network(M, N, seed)


def network_1024x512(seed: int = 0, precision: int = CP_PREC_FP64):
    return network(512, 1024, seed, precision)


def gamma_field(M: int, N: int, seed: int, precision: int = CP_PREC_MIXED, U: float = 0.2, corr: float = 32.0,
                two_five_d: bool = True):
    p = base_params(M, N, 25e-6, U)
    p.alphaMax = 1e9
    p.enableTwoPointFiveD = 1 if two_five_d else 0
    p.Ht = 2e-4
    p.precision = precision
    return p, geom_gamma(M, N, seed, corr)


def gamma_2048(precision: int = CP_PREC_MIXED):
    return gamma_field(2048, 2048, 3, precision)


def fast_mode(p: Params, jacobi_cap: int = 4, pcg_iters: int = 20, cycles: int = 2) -> Params:
    """Inner budgets of the batched / served form (same stop tests, smaller
    caps, FP64-refined pressure, mass residual over the p' columns)."""
    q = p.copy()
    q.precision = CP_PREC_MIXED
    q.maxMomentumIter = jacobi_cap
    q.maxPressureIter = pcg_iters
    q.refineCycles = cycles
    q.pTol = 1e-9
    q.massResidualScope = 1
    return q


def batch_64(seed: int = 4):
    """64 geometries, sizes log-uniform 256x128 .. 1024x512, half networks
    (seeds 100..131), half gamma fields (seeds 200..231), U_in in 0.05-0.3."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(64):
        N = int(round(2 ** rng.uniform(8, 10)))
        N = max(32, (N // 8) * 8)
        M = max(16, (N // 2 // 8) * 8)
        U = float(rng.uniform(0.05, 0.3))
        if k % 2 == 0:
            p, ct = network(M, N, 100 + k // 2, CP_PREC_MIXED, U)
        else:
            p, ct = gamma_field(M, N, 200 + k // 2, CP_PREC_MIXED, U, corr=max(4.0, 32.0 * N / 2048))
        out.append((p, ct))
    return out


def stream_2000(n: int = 2000, seed: int = 5, max_cells: int | None = None):
    """Request list of the served workload: grid sizes log-uniform 128^2 ..
    2048 x 1024, kind uniform over {parallel straight channels, channel
    network, gamma field}, 2.5D on/off at random, U_in in 0.05-0.3. Each
    entry is a small dict; build_request() makes its geometry."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        N = int(round(2 ** rng.uniform(7, 11)))
        M = int(round(2 ** rng.uniform(7, 10)))
        N = max(32, (N // 8) * 8)
        M = max(32, (M // 8) * 8)
        if max_cells and M * N > max_cells:
            s = math.sqrt(max_cells / (M * N))
            M, N = max(32, int(M * s) // 8 * 8), max(32, int(N * s) // 8 * 8)
        kind = int(rng.integers(0, 3))
        twod = bool(rng.integers(0, 2))
        U = float(rng.uniform(0.05, 0.3))
        gseed = int(rng.integers(0, 2**31))
        reqs.append(dict(M=M, N=N, kind=kind, two_five_d=twod, U=U, seed=gseed))
    return reqs


def geom_parallel(M: int, N: int, width: int = 12, wall: int = 4) -> np.ndarray:
    """Parallel straight microchannels across the plate: fluid bands of
    `width` cells separated by `wall`-cell fins, solid border rows."""
    ct = np.ones((M, N))
    r = wall
    while r + width <= M - wall:
        ct[r:r + width, :] = 0.0
        r += width + wall
    if r == wall:  # too thin for one band with fins: one centred channel
        lo = max(1, M // 4)
        ct[lo:M - lo, :] = 0.0
    return ct


def build_request(r: dict):
    M, N = r["M"], r["N"]
    if r["kind"] == 0:
        p = base_params(M, N, 25e-6, r["U"])
        ct = geom_parallel(M, N)
    elif r["kind"] == 1:
        p = network_params(M, N, r["U"])
        ct = geom_network(M, N, r["seed"])
    else:
        p, ct = gamma_field(M, N, r["seed"], CP_PREC_MIXED, r["U"], corr=max(4.0, 32.0 * max(M, N) / 2048))
    p.enableTwoPointFiveD = 1 if r["two_five_d"] else 0
    p.Ht = 2e-4
    p.precision = CP_PREC_MIXED
    return p, ct


def poisson_arrivals(n: int, rate: float, seed: int = 5) -> np.ndarray:
    """Arrival times (s) of n requests of a Poisson process of `rate` per s."""
    rng = np.random.default_rng(seed + 1)
    return np.cumsum(rng.exponential(1.0 / rate, n))
