// cp_thermal.hip — the downstream conjugate-heat-transfer solve on gfx950.
//
// Replaces ThermalSolver3D::solve (ThermalSolver/thermal_solver.cpp:421-604):
// the temperature of a base plate (nz_solid layers) under a channel layer
// (nz_fluid layers) whose fluid carries the SIMPLE velocity field with a
// 4z(1-z) depth profile. The reference assembles a COO matrix on the CPU and
// solves it with AMGCL's IDR(8) + smoothed-aggregation AMG. Here the system
// never leaves the device and never becomes a sparse matrix: it is a 7-point
// stencil stored as seven coefficient planes per layer, and the solver is
// BiCGSTAB preconditioned by one V-cycle of an aggregation multigrid built
// for this geometry (thin in z, wide in x-y):
//   * levels aggregate 2 x 2 cells in the plane and keep every layer
//     (Galerkin: the coarse operator is the sum of its aggregates' rows);
//   * the smoother is z-line Jacobi: each column's tridiagonal z-system
//     is solved exactly (Thomas, factors precomputed per level), the plane
//     couplings taken from the previous iterate; omega 0.8, two sweeps
//     before and after the coarse correction;
//   * the coarsest level (at most 64 columns) runs 50 undamped sweeps in
//     one workgroup, in LDS.
// Dot products reduce per workgroup into fixed slots and then in one
// workgroup in a fixed order, so a solve is bitwise reproducible; the
// BiCGSTAB scalars stay on the device, and the host reads one 64-byte
// record per iteration for the stop test.
//
// Layout: every array of a level is Nz x H x W doubles, cell (i, j, k) at
// (k H + j) W + i — the reference's idx(i, j, k) (thermal_solver.cpp:231)
// at the finest level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/coldplate_defaults.h"
#include "../../include/coldplate_hip.h"
#include "cp_device.h"

namespace {

constexpr int kThBlock = 256;
constexpr int kThMaxNz = 16;        // layers (the z-line kernels are instantiated per count)
constexpr int kThCoarseCols = 64;   // columns of the coarsest level (one wave)
constexpr int kThMaxLevels = 24;
constexpr int kThDotBlocks = 1024;  // workgroups of the reducing kernels (cap)
constexpr int kThCoarseSweeps = 50;
constexpr double kThOmega = 0.8;

// Coefficients of one level: A x at cell c is
//   aP x_c - aW x_{i-1} - aE x_{i+1} - aS x_{j-1} - aN x_{j+1} - aD x_{k-1} - aU x_{k+1}
// (coefficients of absent neighbours are zero), plus the z-line Thomas
// factors: 1/den_k and c_k = aU_k / den_k with den_k = aP_k - aD_k c_{k-1}.
enum ThArr { T_AP = 0, T_AW, T_AE, T_AS, T_AN, T_AD, T_AU, T_INVDEN, T_CFAC, T_NARR };

struct ThLevel {
    int W, H, Nz;
    double* c[T_NARR];
};

// BiCGSTAB scalars kept on the device.
enum ThScalar { S_RHO = 0, S_ALPHA, S_OMEGA, S_BETA, S_RR, S_BB, S_BREAK, S_COUNT = 8 };
enum ThFinOp { F_INIT = 0, F_ALPHA, F_OMEGA, F_RHO };

struct ThAssembleArgs {
    int Nx, Ny, Nz, nzSolid, nzFluid, mode;
    double dx, dy, Hb, Ht, ks, kf, rhoCp, Tin, q0, qk;
    const double* gamma;
    const double* u;
    const double* v;
};

__device__ __host__ inline double thDz(const ThAssembleArgs& a, int k) {
    return k < a.nzSolid ? a.Hb / a.nzSolid : a.Ht / a.nzFluid;
}

// kval (thermal_solver.cpp:260-290): RAMP, harmonic or linear blend of the
// fluid and solid conductivities; every base layer is solid.
__device__ inline double thK(const ThAssembleArgs& a, int cell2d, int k) {
    if (k < a.nzSolid) return a.ks;
    const double g = fmin(1.0, fmax(0.0, a.gamma[cell2d]));
    if (a.mode == 1) return a.kf + (a.ks - a.kf) * (1.0 - g) / (1.0 + a.qk * g);
    if (a.mode == 2) return a.ks * a.kf / ((1.0 - g) * a.kf + g * a.ks + 1e-30);
    return a.ks * (1.0 - g) + a.kf * g;
}

__device__ inline bool thFluid(const ThAssembleArgs& a, int cell2d, int k) {
    return k >= a.nzSolid && a.gamma[cell2d] > 0.01;  // region (:240-244)
}

__device__ inline double thHarm(double x, double y) { return 2 * x * y / (x + y + 1e-30); }

// One finite-volume row per cell (thermal_solver.cpp:440-521), accumulated
// in the reference's order: west (or the inlet), east (or the outlet),
// south, north, below (or the base heat flux), above.
__global__ __launch_bounds__(kThBlock) void k_th_assemble(ThAssembleArgs a, ThLevel L, double* __restrict__ b) {
    const int Nx = a.Nx, Ny = a.Ny, plane = Nx * Ny;
    const int c = blockIdx.x * kThBlock + threadIdx.x;
    if (c >= plane * a.Nz) return;
    const int k = c / plane, c2 = c - k * plane, j = c2 / Nx, i = c2 - j * Nx;
    const double dzk = thDz(a, k), Ax = a.dy * dzk, Ay = a.dx * dzk, Az = a.dx * a.dy;
    const double kP = thK(a, c2, k);
    const double zc = (k - a.nzSolid + 0.5) * (a.Ht / a.nzFluid) / a.Ht;
    const double po = k < a.nzSolid ? 0.0 : 4.0 * zc * (1 - zc);  // pois (:311-315)
    const double gP = k < a.nzSolid ? 0.0 : fmin(1.0, fmax(0.0, a.gamma[c2]));
    const double uP = gP > 0.01 ? a.u[c2] * po : 0.0;
    const double vP = gP > 0.01 ? a.v[c2] * po : 0.0;
    const bool fl = thFluid(a, c2, k);
    double diag = 0, rhs = 0, aW = 0, aE = 0, aS = 0, aN = 0, aD = 0, aU = 0;
    if (i > 0) {
        const double D = thHarm(kP, thK(a, c2 - 1, k)) * Ax / a.dx;
        diag += D;
        aW = D;
        if (fl && thFluid(a, c2 - 1, k)) {
            const double F = a.rhoCp * 0.5 * (uP + a.u[c2 - 1] * po) * Ax;
            diag += fmax(-F, 0.0);
            aW += fmax(F, 0.0);
        }
    } else if (fl) {  // inlet: T = T_inlet at the west face
        const double Dbc = 2.0 * kP * Ax / a.dx, Fbc = a.rhoCp * uP * Ax;
        diag += Dbc;
        rhs += Dbc * a.Tin;
        if (Fbc > 0) rhs += Fbc * a.Tin;
        else diag += -Fbc;
    }
    if (i < Nx - 1) {
        const double D = thHarm(kP, thK(a, c2 + 1, k)) * Ax / a.dx;
        diag += D;
        aE = D;
        if (fl && thFluid(a, c2 + 1, k)) {
            const double F = a.rhoCp * 0.5 * (uP + a.u[c2 + 1] * po) * Ax;
            diag += fmax(F, 0.0);
            aE += fmax(-F, 0.0);
        }
    } else if (fl && uP > 0) {  // convective outlet
        diag += a.rhoCp * uP * Ax;
    }
    if (j > 0) {
        const double D = thHarm(kP, thK(a, c2 - Nx, k)) * Ay / a.dy;
        diag += D;
        aS = D;
        if (fl && thFluid(a, c2 - Nx, k)) {
            const double F = a.rhoCp * 0.5 * (vP + a.v[c2 - Nx] * po) * Ay;
            diag += fmax(-F, 0.0);
            aS += fmax(F, 0.0);
        }
    }
    if (j < Ny - 1) {
        const double D = thHarm(kP, thK(a, c2 + Nx, k)) * Ay / a.dy;
        diag += D;
        aN = D;
        if (fl && thFluid(a, c2 + Nx, k)) {
            const double F = a.rhoCp * 0.5 * (vP + a.v[c2 + Nx] * po) * Ay;
            diag += fmax(F, 0.0);
            aN += fmax(-F, 0.0);
        }
    } else if (fl && vP > 0) {
        diag += a.rhoCp * vP * Ay;
    }
    if (k > 0) {  // series conductance through the layer faces
        const double kN = thK(a, c2, k - 1), dzN = thDz(a, k - 1);
        const double kf = (dzk + dzN) / (dzk / kP + dzN / kN);
        const double D = kf * Az / (0.5 * (dzk + dzN));
        diag += D;
        aD = D;
    } else {
        rhs += a.q0 * Az;  // heat flux into the base
    }
    if (k < a.Nz - 1) {
        const double kN = thK(a, c2, k + 1), dzN = thDz(a, k + 1);
        const double kf = (dzk + dzN) / (dzk / kP + dzN / kN);
        const double D = kf * Az / (0.5 * (dzk + dzN));
        diag += D;
        aU = D;
    }
    L.c[T_AP][c] = diag + 1e-12;
    L.c[T_AW][c] = aW;
    L.c[T_AE][c] = aE;
    L.c[T_AS][c] = aS;
    L.c[T_AN][c] = aN;
    L.c[T_AD][c] = aD;
    L.c[T_AU][c] = aU;
    b[c] = rhs;
}

// Galerkin coarse level of 2 x 2 aggregates: one thread per coarse cell.
__global__ __launch_bounds__(kThBlock) void k_th_coarsen(ThLevel F, ThLevel C) {
    const int planeC = C.W * C.H;
    const int cc = blockIdx.x * kThBlock + threadIdx.x;
    if (cc >= planeC * C.Nz) return;
    const int k = cc / planeC, c2 = cc - k * planeC, J = c2 / C.W, I = c2 - J * C.W;
    double sP = 0, sW = 0, sE = 0, sS = 0, sN = 0, sD = 0, sU = 0;
    for (int bj = 0; bj < 2; ++bj) {
        const int j = 2 * J + bj;
        if (j >= F.H) break;
        for (int bi = 0; bi < 2; ++bi) {
            const int i = 2 * I + bi;
            if (i >= F.W) break;
            const size_t f = (size_t(k) * F.H + j) * F.W + i;
            const double aW = F.c[T_AW][f], aE = F.c[T_AE][f], aS = F.c[T_AS][f], aN = F.c[T_AN][f];
            // couplings inside the aggregate move onto the diagonal (a lone
            // last column or row has none: its east / north coefficient is
            // the zero of the grid boundary)
            sP += F.c[T_AP][f] - (bi == 0 ? aE : aW) - (bj == 0 ? aN : aS);
            if (bi == 0) sW += aW; else sE += aE;
            if (bj == 0) sS += aS; else sN += aN;
            sD += F.c[T_AD][f];
            sU += F.c[T_AU][f];
        }
    }
    C.c[T_AP][cc] = sP;
    C.c[T_AW][cc] = sW;
    C.c[T_AE][cc] = sE;
    C.c[T_AS][cc] = sS;
    C.c[T_AN][cc] = sN;
    C.c[T_AD][cc] = sD;
    C.c[T_AU][cc] = sU;
}

// z-line Thomas factors of a level: one thread per column.
__global__ __launch_bounds__(kThBlock) void k_th_zfactor(ThLevel L) {
    const int plane = L.W * L.H;
    const int c2 = blockIdx.x * kThBlock + threadIdx.x;
    if (c2 >= plane) return;
    double cPrev = 0;
    for (int k = 0; k < L.Nz; ++k) {
        const size_t c = size_t(k) * plane + c2;
        const double den = L.c[T_AP][c] - L.c[T_AD][c] * cPrev;
        const double inv = 1.0 / den;
        cPrev = L.c[T_AU][c] * inv;
        L.c[T_INVDEN][c] = inv;
        L.c[T_CFAC][c] = cPrev;
    }
}

// (A x)_c without the z terms (the plane couplings), for a column kernel.
__device__ inline double thPlaneAx(const ThLevel& L, const double* __restrict__ x, int c, int i, int j) {
    double s = 0;
    if (i > 0) s += L.c[T_AW][c] * x[c - 1];
    if (i < L.W - 1) s += L.c[T_AE][c] * x[c + 1];
    if (j > 0) s += L.c[T_AS][c] * x[c - L.W];
    if (j < L.H - 1) s += L.c[T_AN][c] * x[c + L.W];
    return s;
}

// One z-line Jacobi sweep: xout = xin + omega Z^-1 (b - A xin), Z the
// tridiagonal z-part of A; FROM_ZERO: the iterate is zero (xin unused). NZ
// is the layer count: the column lives in registers and every load of the
// column is independent of the elimination, so all are in flight at once.
template <int NZ, bool FROM_ZERO>
__global__ __launch_bounds__(kThBlock) void k_th_zline(ThLevel L, const double* __restrict__ xin,
                                                       const double* __restrict__ b, double* __restrict__ xout,
                                                       double omega) {
    const int plane = L.W * L.H;
    const int c2 = blockIdx.x * kThBlock + threadIdx.x;
    if (c2 >= plane) return;
    const int j = c2 / L.W, i = c2 - j * L.W;
    double xc[NZ], r[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        const int c = k * plane + c2;
        r[k] = b[c];
        xc[k] = 0.0;
        if (!FROM_ZERO) {
            xc[k] = xin[c];
            r[k] += thPlaneAx(L, xin, c, i, j);
        }
    }
    if (!FROM_ZERO) {
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const int c = k * plane + c2;
            double az = L.c[T_AP][c] * xc[k];
            if (k > 0) az -= L.c[T_AD][c] * xc[k - 1];
            if (k < NZ - 1) az -= L.c[T_AU][c] * xc[k + 1];
            r[k] -= az;
        }
    }
    double d = 0;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        const int c = k * plane + c2;
        d = (r[k] + L.c[T_AD][c] * d) * L.c[T_INVDEN][c];
        r[k] = d;
    }
    double e = 0;
#pragma unroll
    for (int k = NZ - 1; k >= 0; --k) {
        const int c = k * plane + c2;
        e = r[k] + L.c[T_CFAC][c] * e;
        xout[c] = xc[k] + omega * e;
    }
}

// Fine residual b - A x summed over each 2 x 2 aggregate: the coarse
// right-hand side. One thread per coarse cell.
__global__ __launch_bounds__(kThBlock) void k_th_restrict(ThLevel F, const double* __restrict__ x,
                                                          const double* __restrict__ b, int Wc, int Hc,
                                                          double* __restrict__ bc) {
    const int planeC = Wc * Hc, planeF = F.W * F.H;
    const int cc = blockIdx.x * kThBlock + threadIdx.x;
    if (cc >= planeC * F.Nz) return;
    const int k = cc / planeC, c2 = cc - k * planeC, J = c2 / Wc, I = c2 - J * Wc;
    double s = 0;
    for (int bj = 0; bj < 2; ++bj) {
        const int j = 2 * J + bj;
        if (j >= F.H) break;
        for (int bi = 0; bi < 2; ++bi) {
            const int i = 2 * I + bi;
            if (i >= F.W) break;
            const int c = k * planeF + j * F.W + i;
            double ax = F.c[T_AP][c] * x[c] - thPlaneAx(F, x, c, i, j);
            if (k > 0) ax -= F.c[T_AD][c] * x[c - planeF];
            if (k < F.Nz - 1) ax -= F.c[T_AU][c] * x[c + planeF];
            s += b[c] - ax;
        }
    }
    bc[cc] = s;
}

// x += e(aggregate): piecewise-constant prolongation.
__global__ __launch_bounds__(kThBlock) void k_th_prolong(int W, int H, int Nz, const double* __restrict__ ec, int Wc,
                                                         int Hc, double* __restrict__ x) {
    const int plane = W * H;
    const int c = blockIdx.x * kThBlock + threadIdx.x;
    if (c >= plane * Nz) return;
    const int k = c / plane, c2 = c - k * plane, j = c2 / W, i = c2 - j * W;
    x[c] += ec[(size_t(k) * Hc + (j >> 1)) * Wc + (i >> 1)];
}

// The coarsest level: kThCoarseSweeps undamped z-line Jacobi sweeps from
// zero, in one wave, everything in LDS. Undamped, a sweep is
// x <- Z^-1 (b + P x) with P the plane couplings, so with the Thomas
// factors folded in (each row scaled by 1/den_k) a layer costs
//   d_k = b_k/den_k + sum_q (a_q/den_k) x_q + (aD_k/den_k) d_{k-1}
// and the back substitution x_k = d_k + c_k x_{k+1}.
template <int NZ>
__global__ __launch_bounds__(kThCoarseCols) void k_th_coarsest(ThLevel L, const double* __restrict__ b,
                                                               double* __restrict__ x) {
    enum { Q_B = 0, Q_W, Q_E, Q_S, Q_N, Q_L, Q_C, Q_COUNT };
    __shared__ double q[Q_COUNT][NZ * kThCoarseCols];
    __shared__ double xs[2][NZ * kThCoarseCols];
    const int plane = L.W * L.H;
    const int c2 = threadIdx.x;
    const bool on = c2 < plane;
    const int j = on ? c2 / L.W : 0, i = on ? c2 - j * L.W : 0;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        const int s = k * kThCoarseCols + c2;
        xs[0][s] = 0.0;
        if (on) {
            const int c = k * plane + c2;
            const double inv = L.c[T_INVDEN][c];
            q[Q_B][s] = b[c] * inv;
            q[Q_W][s] = i > 0 ? L.c[T_AW][c] * inv : 0.0;
            q[Q_E][s] = i < L.W - 1 ? L.c[T_AE][c] * inv : 0.0;
            q[Q_S][s] = j > 0 ? L.c[T_AS][c] * inv : 0.0;
            q[Q_N][s] = j < L.H - 1 ? L.c[T_AN][c] * inv : 0.0;
            q[Q_L][s] = L.c[T_AD][c] * inv;
            q[Q_C][s] = L.c[T_CFAC][c];
        }
    }
    __syncthreads();
    // neighbour offsets inside a layer; an absent neighbour has a zero
    // coefficient, and its offset points at the cell itself
    const int oW = i > 0 ? -1 : 0, oE = i < L.W - 1 ? 1 : 0;
    const int oS = j > 0 ? -L.W : 0, oN = j < L.H - 1 ? L.W : 0;
    int cur = 0;
    for (int sweep = 0; sweep < kThCoarseSweeps; ++sweep) {
        if (on) {
            const double* xo = xs[cur];
            double* xn = xs[cur ^ 1];
            double d[NZ];
            double dPrev = 0;
#pragma unroll
            for (int k = 0; k < NZ; ++k) {
                const int s = k * kThCoarseCols + c2;
                dPrev = q[Q_B][s] + q[Q_W][s] * xo[s + oW] + q[Q_E][s] * xo[s + oE] + q[Q_S][s] * xo[s + oS] +
                        q[Q_N][s] * xo[s + oN] + q[Q_L][s] * dPrev;
                d[k] = dPrev;
            }
            double e = 0;
#pragma unroll
            for (int k = NZ - 1; k >= 0; --k) {
                const int s = k * kThCoarseCols + c2;
                e = d[k] + q[Q_C][s] * e;
                xn[s] = e;
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    if (on)
#pragma unroll
        for (int k = 0; k < NZ; ++k) x[k * plane + c2] = xs[cur][k * kThCoarseCols + c2];
}

#define TH_FOR_NZ(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
static_assert(kThMaxNz == 16, "TH_FOR_NZ lists every layer count up to kThMaxNz");

// xin == nullptr: a sweep from the zero iterate.
void launchZline(int Nz, int grid, hipStream_t s, const ThLevel& L, const double* xin, const double* b, double* xout,
                 double omega) {
    switch (Nz) {
#define TH_CASE(n)                                                                          \
    case n:                                                                                 \
        if (xin) k_th_zline<n, false><<<grid, kThBlock, 0, s>>>(L, xin, b, xout, omega);   \
        else k_th_zline<n, true><<<grid, kThBlock, 0, s>>>(L, xin, b, xout, omega);        \
        break;
        TH_FOR_NZ(TH_CASE)
#undef TH_CASE
    }
}

void launchCoarsest(int Nz, hipStream_t s, const ThLevel& L, const double* b, double* x) {
    switch (Nz) {
#define TH_CASE(n) \
    case n: k_th_coarsest<n><<<1, kThCoarseCols, 0, s>>>(L, b, x); break;
        TH_FOR_NZ(TH_CASE)
#undef TH_CASE
    }
}

// Workgroup sum of two values into partials[0][blockIdx], partials[1][blockIdx]
// (fixed tree: the result does not depend on timing).
__device__ inline void thBlockSum2(double a, double b, double* __restrict__ partials) {
    __shared__ double sa[kThBlock], sb[kThBlock];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int w = kThBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = sa[0];
        partials[gridDim.x + blockIdx.x] = sb[0];
    }
}

__device__ inline double thAx(const ThLevel& L, const double* __restrict__ x, int c, int plane) {
    const int k = c / plane, c2 = c - k * plane, j = c2 / L.W, i = c2 - j * L.W;
    double ax = L.c[T_AP][c] * x[c] - thPlaneAx(L, x, c, i, j);
    if (k > 0) ax -= L.c[T_AD][c] * x[c - plane];
    if (k < L.Nz - 1) ax -= L.c[T_AU][c] * x[c + plane];
    return ax;
}

// r = b - A x; rhat = p = r; v = 0; partials of (r, r) and (b, b).
__global__ __launch_bounds__(kThBlock) void k_th_init(ThLevel L, const double* __restrict__ x,
                                                      const double* __restrict__ b, double* __restrict__ r,
                                                      double* __restrict__ rhat, double* __restrict__ p,
                                                      double* __restrict__ v, double* __restrict__ partials) {
    const int plane = L.W * L.H;
    const int n = plane * L.Nz;
    double rr = 0, bb = 0;
    for (int c = blockIdx.x * kThBlock + threadIdx.x; c < n; c += gridDim.x * kThBlock) {
        const double bc = b[c], rc = bc - thAx(L, x, c, plane);
        r[c] = rc;
        rhat[c] = rc;
        p[c] = rc;
        v[c] = 0.0;
        rr += rc * rc;
        bb += bc * bc;
    }
    thBlockSum2(rr, bb, partials);
}

// y = A x; partials of (w, y) and (y, y).
__global__ __launch_bounds__(kThBlock) void k_th_matvec(ThLevel L, const double* __restrict__ x,
                                                        double* __restrict__ y, const double* __restrict__ w,
                                                        double* __restrict__ partials) {
    const int plane = L.W * L.H;
    const int n = plane * L.Nz;
    double wy = 0, yy = 0;
    for (int c = blockIdx.x * kThBlock + threadIdx.x; c < n; c += gridDim.x * kThBlock) {
        const double yc = thAx(L, x, c, plane);
        y[c] = yc;
        wy += w[c] * yc;
        yy += yc * yc;
    }
    thBlockSum2(wy, yy, partials);
}

// p = r + beta (p - omega v)
__global__ __launch_bounds__(kThBlock) void k_th_pupdate(int n, const double* __restrict__ sc,
                                                         const double* __restrict__ r, const double* __restrict__ v,
                                                         double* __restrict__ p) {
    const double beta = sc[S_BETA], omega = sc[S_OMEGA];
    for (int c = blockIdx.x * kThBlock + threadIdx.x; c < n; c += gridDim.x * kThBlock)
        p[c] = r[c] + beta * (p[c] - omega * v[c]);
}

// s = r - alpha v
__global__ __launch_bounds__(kThBlock) void k_th_supdate(int n, const double* __restrict__ sc,
                                                         const double* __restrict__ r, const double* __restrict__ v,
                                                         double* __restrict__ s) {
    const double alpha = sc[S_ALPHA];
    for (int c = blockIdx.x * kThBlock + threadIdx.x; c < n; c += gridDim.x * kThBlock)
        s[c] = r[c] - alpha * v[c];
}

// x += alpha phat + omega shat; r = s - omega t; partials of (rhat, r), (r, r).
__global__ __launch_bounds__(kThBlock) void k_th_xrupdate(int n, const double* __restrict__ sc,
                                                          const double* __restrict__ phat,
                                                          const double* __restrict__ shat,
                                                          const double* __restrict__ s, const double* __restrict__ t,
                                                          const double* __restrict__ rhat, double* __restrict__ x,
                                                          double* __restrict__ r, double* __restrict__ partials) {
    const double alpha = sc[S_ALPHA], omega = sc[S_OMEGA];
    double hr = 0, rr = 0;
    for (int c = blockIdx.x * kThBlock + threadIdx.x; c < n; c += gridDim.x * kThBlock) {
        x[c] += alpha * phat[c] + omega * shat[c];
        const double rc = s[c] - omega * t[c];
        r[c] = rc;
        hr += rhat[c] * rc;
        rr += rc * rc;
    }
    thBlockSum2(hr, rr, partials);
}

// Sum the partials of two dot products in a fixed order and update the
// BiCGSTAB scalars. One workgroup.
__global__ __launch_bounds__(kThBlock) void k_th_fin(const double* __restrict__ partials, int nblk, int op,
                                                     double* __restrict__ sc) {
    __shared__ double sa[kThBlock], sb[kThBlock];
    double a = 0, b = 0;
    for (int q = threadIdx.x; q < nblk; q += kThBlock) {
        a += partials[q];
        b += partials[nblk + q];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int w = kThBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double d0 = sa[0], d1 = sb[0];
    switch (op) {
        case F_INIT:  // d0 = (r, r) = (rhat, r), d1 = (b, b)
            sc[S_RR] = d0;
            sc[S_BB] = d1;
            sc[S_RHO] = d0;
            sc[S_ALPHA] = 1.0;
            sc[S_OMEGA] = 1.0;
            sc[S_BETA] = 0.0;
            sc[S_BREAK] = 0.0;
            break;
        case F_ALPHA:  // d0 = (rhat, v)
            if (d0 != 0.0 && isfinite(d0)) {
                sc[S_ALPHA] = sc[S_RHO] / d0;
            } else {
                sc[S_ALPHA] = 0.0;
                sc[S_BREAK] = 1.0;
            }
            break;
        case F_OMEGA:  // d0 = (t, s), d1 = (t, t)
            sc[S_OMEGA] = d1 > 0.0 ? d0 / d1 : 0.0;
            break;
        case F_RHO: {  // d0 = (rhat, r), d1 = (r, r)
            const double rho = sc[S_RHO], omega = sc[S_OMEGA];
            sc[S_RR] = d1;
            if (rho != 0.0 && omega != 0.0) {
                sc[S_BETA] = (d0 / rho) * (sc[S_ALPHA] / omega);
            } else {
                sc[S_BETA] = 0.0;
                sc[S_BREAK] = 1.0;
            }
            sc[S_RHO] = d0;
            if (d0 == 0.0) sc[S_BREAK] = 1.0;
            break;
        }
    }
}

int gridOf(size_t n) { return int(std::max<size_t>(1, (n + kThBlock - 1) / kThBlock)); }
int dotGrid(size_t n) { return std::min(gridOf(n), kThDotBlocks); }

struct ThDev {
    std::vector<void*> allocs;
    std::vector<hipEvent_t> events;
    hipStream_t stream = nullptr;
    double* hostSc = nullptr;
    ~ThDev() {  // teardown: errors here have no caller left to report to
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (void* p : allocs) (void)hipFree(p);
        if (hostSc) (void)hipHostFree(hostSc);
        if (stream) (void)hipStreamDestroy(stream);
    }
    double* alloc(size_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) return nullptr;
        allocs.push_back(p);
        return static_cast<double*>(p);
    }
};

#define TH_TRY(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return cpdev::reportError(CP_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// One V-cycle: x[0] = M^-1 b0 (x[0] and b0 are the caller's arrays).
struct ThHierarchy {
    std::vector<ThLevel> lv;
    std::vector<double*> x, b, y;  // per level (level 0: x, b set per call; y scratch)
    // timing: event pairs around the finest level's sweeps from a nonzero
    // iterate (three per V-cycle); evUsed pairs recorded since the last read
    hipEvent_t* ev = nullptr;
    int evCap = 0, evUsed = 0;

    void fineSweep(hipStream_t s, const ThLevel& F, const double* xin, const double* bb, double* xout) {
        const bool timed = ev && evUsed < evCap;
        if (timed) (void)hipEventRecord(ev[2 * evUsed], s);
        launchZline(F.Nz, gridOf(size_t(F.W) * F.H), s, F, xin, bb, xout, kThOmega);
        if (timed) (void)hipEventRecord(ev[2 * evUsed++ + 1], s);
    }

    // Sum of the recorded sweeps' durations (ms), after a stream sync.
    double readSweeps(int* count) {
        double ms = 0;
        for (int q = 0; q < evUsed; ++q) {
            float t = 0;
            if (hipEventElapsedTime(&t, ev[2 * q], ev[2 * q + 1]) == hipSuccess) ms += t;
        }
        *count += evUsed;
        evUsed = 0;
        return ms;
    }

    hipError_t vcycle(hipStream_t s, const double* b0, double* x0) {
        b[0] = const_cast<double*>(b0);
        x[0] = x0;
        const int L = int(lv.size());
        for (int l = 0; l < L - 1; ++l) {
            const ThLevel& F = lv[l];
            const int cols = F.W * F.H;
            launchZline(F.Nz, gridOf(cols), s, F, nullptr, b[l], y[l], kThOmega);
            if (l == 0) fineSweep(s, F, y[l], b[l], x[l]);
            else launchZline(F.Nz, gridOf(cols), s, F, y[l], b[l], x[l], kThOmega);
            const ThLevel& C = lv[l + 1];
            k_th_restrict<<<gridOf(size_t(C.W) * C.H * C.Nz), kThBlock, 0, s>>>(F, x[l], b[l], C.W, C.H, b[l + 1]);
        }
        launchCoarsest(lv[L - 1].Nz, s, lv[L - 1], b[L - 1], x[L - 1]);
        for (int l = L - 2; l >= 0; --l) {
            const ThLevel& F = lv[l];
            const ThLevel& C = lv[l + 1];
            const int cols = F.W * F.H;
            k_th_prolong<<<gridOf(size_t(cols) * F.Nz), kThBlock, 0, s>>>(F.W, F.H, F.Nz, x[l + 1], C.W, C.H, x[l]);
            if (l == 0) {
                fineSweep(s, F, x[l], b[l], y[l]);
                fineSweep(s, F, y[l], b[l], x[l]);
            } else {
                launchZline(F.Nz, gridOf(cols), s, F, x[l], b[l], y[l], kThOmega);
                launchZline(F.Nz, gridOf(cols), s, F, y[l], b[l], x[l], kThOmega);
            }
        }
        return hipGetLastError();
    }
};

void hostStats(const cp_thermal_params& p, const double* gamma, const double* u, const double* T,
               cp_thermal_stats* s);

}  // namespace

extern "C" {

int cp_thermal_params_default(cp_thermal_params* p) {
    if (!p) return cpdev::reportError(CP_E_ARG, "null params");
    cp_fill_thermal_defaults(p);
    return CP_OK;
}

int cp_thermal_solve(const cp_thermal_params* prm, const double* gamma, const double* u, const double* v, double* T,
                     cp_thermal_stats* st) {
    if (!prm || !gamma || !u || !v || !T || !st) return cpdev::reportError(CP_E_ARG, "null argument");
    const cp_thermal_params p = *prm;
    if (p.Nx < 1 || p.Ny < 1) return cpdev::reportError(CP_E_ARG, "grid must be at least 1 x 1 (Nx=%d Ny=%d)", p.Nx, p.Ny);
    if (p.nz_solid < 1 || p.nz_fluid < 1 || p.nz_solid + p.nz_fluid > kThMaxNz)
        return cpdev::reportError(CP_E_ARG, "layers: nz_solid, nz_fluid >= 1 and nz_solid + nz_fluid <= %d", kThMaxNz);
    if (!(p.dx > 0) || !(p.dy > 0) || !(p.H_b > 0) || !(p.H_t > 0))
        return cpdev::reportError(CP_E_ARG, "dx, dy, H_b, H_t must be > 0");
    if (!(p.k_s > 0) || !(p.k_f > 0) || !(p.rho > 0) || !(p.Cp > 0))
        return cpdev::reportError(CP_E_ARG, "k_s, k_f, rho, Cp must be > 0");
    if (p.k_interp_mode < 0 || p.k_interp_mode > 2)
        return cpdev::reportError(CP_E_ARG, "unknown k_interp_mode %d (0 linear, 1 RAMP, 2 harmonic)", p.k_interp_mode);
    if (!(p.rtol > 0) || p.maxiter < 1) return cpdev::reportError(CP_E_ARG, "rtol must be > 0 and maxiter >= 1");
    const size_t plane = size_t(p.Nx) * p.Ny, Nz = size_t(p.nz_solid + p.nz_fluid), n = plane * Nz;
    if (n > (size_t(1) << 30)) return cpdev::reportError(CP_E_ARG, "grid too large (%zu cells, at most 2^30)", n);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return cpdev::reportError(CP_E_NODEV, "no HIP device");
    if (!cpdev::currentDeviceIsGfx950())
        return cpdev::reportError(CP_E_NODEV, "the current HIP device is not a gfx950 (MI355X) device");

    ThDev dev;
    TH_TRY(hipStreamCreateWithFlags(&dev.stream, hipStreamNonBlocking));
    TH_TRY(hipHostMalloc(reinterpret_cast<void**>(&dev.hostSc), S_COUNT * sizeof(double)));
    hipStream_t s = dev.stream;

    // the hierarchy
    ThHierarchy h;
    {
        int W = p.Nx, H = p.Ny;
        for (;;) {
            ThLevel L{};
            L.W = W;
            L.H = H;
            L.Nz = int(Nz);
            h.lv.push_back(L);
            if (size_t(W) * H <= size_t(kThCoarseCols) || int(h.lv.size()) == kThMaxLevels) break;
            W = (W + 1) / 2;
            H = (H + 1) / 2;
        }
        if (size_t(h.lv.back().W) * h.lv.back().H > size_t(kThCoarseCols))
            return cpdev::reportError(CP_E_ARG, "grid too elongated for %d multigrid levels", kThMaxLevels);
    }
    const int nl = int(h.lv.size());
    h.x.assign(nl, nullptr);
    h.b.assign(nl, nullptr);
    h.y.assign(nl, nullptr);
    for (int l = 0; l < nl; ++l) {
        ThLevel& L = h.lv[l];
        const size_t nlv = size_t(L.W) * L.H * L.Nz;
        for (int a = 0; a < T_NARR; ++a)
            if (!(L.c[a] = dev.alloc(nlv))) return cpdev::reportError(CP_E_NOMEM, "hipMalloc of the thermal hierarchy");
        if (!(h.y[l] = dev.alloc(nlv))) return cpdev::reportError(CP_E_NOMEM, "hipMalloc of the thermal hierarchy");
        if (l > 0 && (!(h.x[l] = dev.alloc(nlv)) || !(h.b[l] = dev.alloc(nlv))))
            return cpdev::reportError(CP_E_NOMEM, "hipMalloc of the thermal hierarchy");
    }
    // Krylov vectors and inputs
    enum { V_X, V_B, V_R, V_RHAT, V_P, V_V, V_PHAT, V_S, V_SHAT, V_T, V_COUNT };
    double* vec[V_COUNT];
    for (int a = 0; a < V_COUNT; ++a)
        if (!(vec[a] = dev.alloc(n))) return cpdev::reportError(CP_E_NOMEM, "hipMalloc of %zu cells", n);
    double *dGamma = dev.alloc(plane), *dU = dev.alloc(plane), *dV = dev.alloc(plane);
    const int nblk = dotGrid(n);
    double* partials = dev.alloc(2 * size_t(nblk));
    double* sc = dev.alloc(S_COUNT);
    if (!dGamma || !dU || !dV || !partials || !sc) return cpdev::reportError(CP_E_NOMEM, "hipMalloc of the inputs");
    TH_TRY(hipMemcpyAsync(dGamma, gamma, plane * sizeof(double), hipMemcpyHostToDevice, s));
    TH_TRY(hipMemcpyAsync(dU, u, plane * sizeof(double), hipMemcpyHostToDevice, s));
    TH_TRY(hipMemcpyAsync(dV, v, plane * sizeof(double), hipMemcpyHostToDevice, s));

    // timing events: the solve's start and end, and six sweep pairs (the
    // finest level's timed sweeps of one BiCGSTAB iteration)
    hipEvent_t evSolve[2] = {nullptr, nullptr};
    double sweepMs = 0;
    int sweepCount = 0;
    if (p.timing) {
        dev.events.resize(2 + 12, nullptr);
        for (auto& e : dev.events) TH_TRY(hipEventCreate(&e));
        evSolve[0] = dev.events[0];
        evSolve[1] = dev.events[1];
        h.ev = dev.events.data() + 2;
        h.evCap = 6;
        TH_TRY(hipEventRecord(evSolve[0], s));
    }

    // assemble, coarsen, factor
    ThAssembleArgs a{};
    a.Nx = p.Nx;
    a.Ny = p.Ny;
    a.Nz = int(Nz);
    a.nzSolid = p.nz_solid;
    a.nzFluid = p.nz_fluid;
    a.mode = p.k_interp_mode;
    a.dx = p.dx;
    a.dy = p.dy;
    a.Hb = p.H_b;
    a.Ht = p.H_t;
    a.ks = p.k_s;
    a.kf = p.k_f;
    a.rhoCp = p.rho * p.Cp;
    a.Tin = p.T_inlet;
    a.q0 = p.q0;
    a.qk = p.qk;
    a.gamma = dGamma;
    a.u = dU;
    a.v = dV;
    k_th_assemble<<<gridOf(n), kThBlock, 0, s>>>(a, h.lv[0], vec[V_B]);
    for (int l = 0; l + 1 < nl; ++l)
        k_th_coarsen<<<gridOf(size_t(h.lv[l + 1].W) * h.lv[l + 1].H * Nz), kThBlock, 0, s>>>(h.lv[l], h.lv[l + 1]);
    for (int l = 0; l < nl; ++l) k_th_zfactor<<<gridOf(size_t(h.lv[l].W) * h.lv[l].H), kThBlock, 0, s>>>(h.lv[l]);
    TH_TRY(hipGetLastError());

    // T = T_inlet everywhere (thermal_solver.cpp:427)
    {
        std::vector<double> t0(n, p.T_inlet);
        TH_TRY(hipMemcpyAsync(vec[V_X], t0.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
        TH_TRY(hipStreamSynchronize(s));
    }

    const ThLevel& L0 = h.lv[0];
    auto readScalars = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(dev.hostSc, sc, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        return e;
    };
    auto restart = [&]() -> hipError_t {
        k_th_init<<<nblk, kThBlock, 0, s>>>(L0, vec[V_X], vec[V_B], vec[V_R], vec[V_RHAT], vec[V_P], vec[V_V], partials);
        k_th_fin<<<1, kThBlock, 0, s>>>(partials, nblk, F_INIT, sc);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? readScalars() : e;
    };
    TH_TRY(restart());
    const double bb = dev.hostSc[S_BB];
    double relres = bb > 0 ? std::sqrt(dev.hostSc[S_RR] / bb) : 0.0;
    int iters = 0, converged = relres < p.rtol ? 1 : 0;
    bool fresh = true;  // p = r and v = 0 were just set by a restart
    while (!converged && iters < p.maxiter) {
        if (!fresh) k_th_pupdate<<<nblk, kThBlock, 0, s>>>(int(n), sc, vec[V_R], vec[V_V], vec[V_P]);
        fresh = false;
        TH_TRY(h.vcycle(s, vec[V_P], vec[V_PHAT]));
        k_th_matvec<<<nblk, kThBlock, 0, s>>>(L0, vec[V_PHAT], vec[V_V], vec[V_RHAT], partials);
        k_th_fin<<<1, kThBlock, 0, s>>>(partials, nblk, F_ALPHA, sc);
        k_th_supdate<<<nblk, kThBlock, 0, s>>>(int(n), sc, vec[V_R], vec[V_V], vec[V_S]);
        TH_TRY(h.vcycle(s, vec[V_S], vec[V_SHAT]));
        // (t, s) and (t, t): the weight of the first dot is s
        k_th_matvec<<<nblk, kThBlock, 0, s>>>(L0, vec[V_SHAT], vec[V_T], vec[V_S], partials);
        k_th_fin<<<1, kThBlock, 0, s>>>(partials, nblk, F_OMEGA, sc);
        k_th_xrupdate<<<nblk, kThBlock, 0, s>>>(int(n), sc, vec[V_PHAT], vec[V_SHAT], vec[V_S], vec[V_T], vec[V_RHAT],
                                                vec[V_X], vec[V_R], partials);
        k_th_fin<<<1, kThBlock, 0, s>>>(partials, nblk, F_RHO, sc);
        TH_TRY(hipGetLastError());
        TH_TRY(readScalars());
        if (p.timing) sweepMs += h.readSweeps(&sweepCount);
        ++iters;
        const double rr = dev.hostSc[S_RR];
        if (!std::isfinite(rr)) return cpdev::reportError(CP_E_DIVERGED, "thermal residual is not finite at iteration %d", iters);
        relres = std::sqrt(rr / bb);
        if (relres < p.rtol || dev.hostSc[S_BREAK] != 0.0) {
            // the recurrence residual met the tolerance (or BiCGSTAB broke
            // down): recompute the true residual and start over from it
            TH_TRY(restart());
            relres = std::sqrt(dev.hostSc[S_RR] / bb);
            if (!std::isfinite(relres))
                return cpdev::reportError(CP_E_DIVERGED, "thermal residual is not finite at iteration %d", iters);
            converged = relres < p.rtol ? 1 : 0;
            fresh = true;
        }
    }
    if (!converged) {  // the true residual of the returned field
        TH_TRY(restart());
        relres = bb > 0 ? std::sqrt(dev.hostSc[S_RR] / bb) : 0.0;
        if (!std::isfinite(relres)) return cpdev::reportError(CP_E_DIVERGED, "thermal residual is not finite");
    }
    if (p.timing) TH_TRY(hipEventRecord(evSolve[1], s));
    TH_TRY(hipMemcpyAsync(T, vec[V_X], n * sizeof(double), hipMemcpyDeviceToHost, s));
    TH_TRY(hipStreamSynchronize(s));
    std::memset(st, 0, sizeof(*st));
    st->iters = iters;
    st->converged = converged;
    st->relres = relres;
    if (p.timing) {
        float ms = 0;
        TH_TRY(hipEventElapsedTime(&ms, evSolve[0], evSolve[1]));
        st->device_ms = ms;
        st->sweeps = sweepCount;
        st->sweep_ms = sweepCount ? sweepMs / sweepCount : 0.0;
    }
    hostStats(p, gamma, u, T, st);
    return CP_OK;
}

}  // extern "C"

namespace {

// Temperature statistics and energy balance of thermal_solver.cpp:556-604
// and the volume-weighted averages of thermal_output.cpp:135-205, over the
// field already copied back.
void hostStats(const cp_thermal_params& p, const double* gamma, const double* u, const double* T,
               cp_thermal_stats* s) {
    const int Nx = p.Nx, Ny = p.Ny, Nz = p.nz_solid + p.nz_fluid;
    const size_t plane = size_t(Nx) * Ny;
    auto dz = [&](int k) { return k < p.nz_solid ? p.H_b / p.nz_solid : p.H_t / p.nz_fluid; };
    auto pois = [&](int k) {
        const double z = (k - p.nz_solid + 0.5) * (p.H_t / p.nz_fluid) / p.H_t;
        return k < p.nz_solid ? 0.0 : 4.0 * z * (1 - z);
    };
    s->T_base_min = s->T_fin_min = s->T_fluid_min = 1e30;
    s->T_base_max = s->T_fin_max = s->T_fluid_max = -1e30;
    double tvB = 0, vB = 0, tvF = 0, vF = 0, tvS = 0, vS = 0, tvG = 0, vG = 0;
    for (int k = 0; k < Nz; ++k) {
        const double V = p.dx * p.dy * dz(k);
        for (size_t c2 = 0; c2 < plane; ++c2) {
            const double t = T[size_t(k) * plane + c2];
            tvG += t * V;
            vG += V;
            if (k < p.nz_solid) {
                s->T_base_min = std::min(s->T_base_min, t);
                s->T_base_max = std::max(s->T_base_max, t);
                tvB += t * V;
                vB += V;
            } else if (gamma[c2] > 0.01) {
                s->T_fluid_min = std::min(s->T_fluid_min, t);
                s->T_fluid_max = std::max(s->T_fluid_max, t);
                tvF += t * V;
                vF += V;
            } else {
                s->T_fin_min = std::min(s->T_fin_min, t);
                s->T_fin_max = std::max(s->T_fin_max, t);
                tvS += t * V;
                vS += V;
            }
        }
    }
    s->T_avg_base = vB > 0 ? tvB / vB : 0.0;
    s->T_avg_fluid = vF > 0 ? tvF / vF : 0.0;
    s->T_avg_solid = vS > 0 ? tvS / vS : 0.0;
    s->T_avg_global = vG > 0 ? tvG / vG : 0.0;
    double Qin = 0, mdot = 0, mdotT = 0;
    for (size_t c2 = 0; c2 < plane; ++c2) Qin += p.q0 * p.dx * p.dy;
    for (int k = p.nz_solid; k < Nz; ++k) {
        const double A = p.dy * dz(k), po = pois(k);
        for (int j = 0; j < Ny; ++j) {
            const size_t c2 = size_t(j) * Nx + Nx - 1;
            if (gamma[c2] > 0.01) {
                const double uu = u[c2] * po;
                if (uu > 0) {
                    const double m = p.rho * uu * A;
                    mdot += m;
                    mdotT += m * T[size_t(k) * plane + c2];
                }
            }
        }
    }
    s->T_out = mdot > 0 ? mdotT / mdot : p.T_inlet;
    s->Q_in = Qin;
    s->Q_out = mdot * p.Cp * (s->T_out - p.T_inlet);
    s->imbalance_pct = 100 * (Qin - s->Q_out) / (Qin + 1e-30);
}

}  // namespace
