// simple_oracle.hpp — CPU restatement of the reference SIMPLE hot path.
//
// TEST INFRASTRUCTURE ONLY. This is the oracle the GPU backend is checked
// against, and (instantiated with float) the reference-equivalent CPU
// baseline. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg may load it. The product path never calls it.
//
// Scalar = double : the FP64 oracle.
// Scalar = float  : same arithmetic as the reference (all-float fields,
//                   SIMPLE.h:399-440), used as the CPU baseline.
//
// Shapes and index ranges are the reference's (row-major here; the
// reference's Eigen storage is column-major, which only changes the memory
// order, not the values):
//   u, uStar, uOld, dE : (M+1) x N      v, vStar, vOld, dN : M x (N+1)
//   p, pStar           : (M+1) x (N+1)  cellType, gamma, alpha : M x N
//
// Reductions: max-reductions are order independent. Sum-reductions use a
// fixed blocking (contiguous blocks of kBlock rows summed serially, then the
// block partials summed serially), so results do not depend on the OpenMP
// thread count. For n <= kBlock the order is exactly sequential, as the
// reference's solvers run with one thread.
#pragma once

#include "../include/coldplate_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {

constexpr int kBlock = 1 << 16;  // sums are sequential up to 65536 rows

template <typename T>
T blockedSum(int n, const T* a, const T* b) {
    const int nb = (n + kBlock - 1) / kBlock;
    std::vector<T> part(std::max(nb, 1), T(0));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; ++k) {
        const int lo = k * kBlock, hi = std::min(n, lo + kBlock);
        T s = T(0);
        for (int i = lo; i < hi; ++i) s += a[i] * b[i];
        part[k] = s;
    }
    T s = T(0);
    for (int k = 0; k < nb; ++k) s += part[k];
    return s;
}

// CSR system with the per-row offsets of the diagonal and the four
// neighbours (momentum_solver.cpp:36-42, 79-159). Columns of a row are
// sorted ascending, which for the row-major DOF numbering is the order
// S, W, P, E, N.
template <typename T>
struct CSR {
    int n = 0;
    std::vector<int> row_ptr, col_idx;
    std::vector<T> values, rhs, sol;
    std::vector<int> diagOff, eOff, wOff, nOff, sOff;
    std::vector<int> ri, rj;  // DOF -> grid (i, j)
};

// y = A x (linalg.h:35-51, pressure_solver.cpp:108-118).
template <typename T>
void spmv(const CSR<T>& A, const T* x, T* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A.n; ++i) {
        T s = T(0);
        for (int k = A.row_ptr[i]; k < A.row_ptr[i + 1]; ++k) s += A.values[k] * x[A.col_idx[k]];
        y[i] = s;
    }
}

// Point Jacobi with double buffering (Utilities/solvers/cpu/jacobi_solver.cpp:20-75):
//   r = b - A x ; x' = x + r / A_ii ; stop after the sweep in which
//   sum(r^2) < tol^2 n ; at most maxIter sweeps. Returns the sweep count.
template <typename T>
int jacobi(int n, const int* row_ptr, const int* col_idx, const T* values, const T* b, T* x,
           const int* diagOff, T tol, int maxIter) {
    if (n == 0) return 0;
    std::vector<T> buf(n);
    T* cur = x;
    T* nxt = buf.data();
    const T tolSq = tol * tol;
    const int nb = (n + kBlock - 1) / kBlock;
    std::vector<T> part(nb);
    int iter = 0;
    for (; iter < maxIter; ++iter) {
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nb; ++k) {
            const int lo = k * kBlock, hi = std::min(n, lo + kBlock);
            T s = T(0);
            for (int i = lo; i < hi; ++i) {
                T Ax = T(0);
                for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q) Ax += values[q] * cur[col_idx[q]];
                const T r = b[i] - Ax;
                s += r * r;
                nxt[i] = cur[i] + r / values[diagOff[i]];
            }
            part[k] = s;
        }
        T total = T(0);
        for (int k = 0; k < nb; ++k) total += part[k];
        std::swap(cur, nxt);
        if (total < tolSq * T(n)) {  // jacobi_solver.cpp:63
            iter++;
            break;
        }
    }
    if (cur != x) std::copy(cur, cur + n, x);
    return iter;
}

// Jacobi-preconditioned CG, x0 given (pressure_solver.cpp:121-184, the copy
// that solvePressureSystem calls; algorithmically equal to
// Utilities/solvers/cpu/pcg_solver.cpp:18-92). Stops on |pAp| < 1e-20, on the
// absolute ||r||_2 < tol, or after maxIter iterations; returns the loop
// index at the break (maxIter when the cap is hit).
template <typename T>
int pcg(const CSR<T>& A, const T* b, T* x, T tol, int maxIter) {
    const int n = A.n;
    if (n == 0) return 0;
    std::vector<T> r(n), z(n), p(n), Ap(n), diag(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        T d = T(1);
        for (int k = A.row_ptr[i]; k < A.row_ptr[i + 1]; ++k)
            if (A.col_idx[k] == i) { d = A.values[k]; break; }
        diag[i] = d;
    }
    spmv(A, x, Ap.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    auto applyJac = [&]() {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) z[i] = (std::abs(diag[i]) > T(1e-20)) ? r[i] / diag[i] : r[i];
    };
    applyJac();
    p = z;
    T rz = blockedSum(n, r.data(), z.data());
    int iter = 0;
    for (; iter < maxIter; ++iter) {
        spmv(A, p.data(), Ap.data());
        const T pAp = blockedSum(n, p.data(), Ap.data());
        if (std::abs(pAp) < T(1e-20)) break;
        const T alpha = rz / pAp;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        const T err = blockedSum(n, r.data(), r.data());
        if (std::sqrt(err) < tol) break;
        applyJac();
        const T rzPrev = rz;
        rz = blockedSum(n, r.data(), z.data());
        const T beta = rz / rzPrev;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    return iter;
}

// ---------------------------------------------------------------------------
// Aggregation-multigrid preconditioner of the pressure-correction system
// (cp_params.pressurePrecond = 1; no counterpart in the reference, whose CPU
// PCG is Jacobi-preconditioned and whose AMGCL backends are not restated).
// The unknowns form the H x W grid of p' DOFs (I = i-1, J = j-1); the
// operator is 5-point: A x = aP x - cE x_E - cE_W x_W - cN x_N - cN_S x_S,
// cE(I,J) coupling (I,J)-(I,J+1) and cN(I,J) coupling (I,J)-(I+1,J).
// Levels: 2x2 piecewise-constant aggregates, Galerkin coarse operator
// A_c = P^T A P (5-point again); coarsening stops once a level has at most
// kCoarseMax unknowns (at least one coarse level). V-cycle: two
// damped-Jacobi sweeps from zero, restriction of the residual (sum over the
// aggregate), coarse correction scaled by `scale`, two damped-Jacobi sweeps;
// the coarsest level is solved with its dense inverse (Gauss-Jordan in
// double, no pivoting: the coarse matrices are symmetric positive definite).
// The device kernels follow the same operation order.
// ---------------------------------------------------------------------------
template <typename T>
struct AggMG {
    static constexpr int kMaxLevels = 12;
    static constexpr int kCoarseMax = 64;
    struct Level {
        int H = 0, W = 0;
        std::vector<T> aP, cE, cN, b, x, z;
        size_t at(int I, int J) const { return size_t(I) * W + J; }
    };
    std::vector<Level> lv;
    std::vector<double> inv;  // coarsest level, n x n row-major
    T omega = T(0.8), scale = T(1.8);

    // A x at (I,J), terms in the CSR column order S, W, P, E, N.
    static T ax(const Level& L, const T* x, int I, int J) {
        const size_t k = L.at(I, J);
        T s = T(0);
        if (I > 0) s += -L.cN[k - L.W] * x[k - L.W];
        if (J > 0) s += -L.cE[k - 1] * x[k - 1];
        s += L.aP[k] * x[k];
        if (J + 1 < L.W) s += -L.cE[k] * x[k + 1];
        if (I + 1 < L.H) s += -L.cN[k] * x[k + L.W];
        return s;
    }

    static void coarsen(const Level& f, Level& c) {
        c.H = (f.H + 1) / 2;
        c.W = (f.W + 1) / 2;
        const size_t n = size_t(c.H) * c.W;
        c.aP.assign(n, T(0)); c.cE.assign(n, T(0)); c.cN.assign(n, T(0));
        c.b.assign(n, T(0)); c.x.assign(n, T(0)); c.z.assign(n, T(0));
        auto in = [&](int I, int J) { return I < f.H && J < f.W; };
        auto g = [&](const std::vector<T>& a, int I, int J) { return in(I, J) ? double(a[f.at(I, J)]) : 0.0; };
        // coupling arrays restricted to edges whose both ends are unknowns
        auto eE = [&](int I, int J) { return (in(I, J) && J + 1 < f.W) ? double(f.cE[f.at(I, J)]) : 0.0; };
        auto eN = [&](int I, int J) { return (in(I, J) && I + 1 < f.H) ? double(f.cN[f.at(I, J)]) : 0.0; };
        for (int I = 0; I < c.H; ++I)
            for (int J = 0; J < c.W; ++J) {
                const int i0 = 2 * I, j0 = 2 * J;
                double acc = 0.0;
                acc += g(f.aP, i0, j0);
                acc += g(f.aP, i0, j0 + 1);
                acc += g(f.aP, i0 + 1, j0);
                acc += g(f.aP, i0 + 1, j0 + 1);
                double internal = 0.0;
                internal += eE(i0, j0);
                internal += eE(i0 + 1, j0);
                internal += eN(i0, j0);
                internal += eN(i0, j0 + 1);
                const size_t k = c.at(I, J);
                c.aP[k] = T(acc - 2.0 * internal);
                c.cE[k] = T(eE(i0, j0 + 1) + eE(i0 + 1, j0 + 1));
                c.cN[k] = T(eN(i0 + 1, j0) + eN(i0 + 1, j0 + 1));
            }
    }

    // Dense inverse of the coarsest operator by Gauss-Jordan elimination.
    void factor() {
        const Level& L = lv.back();
        const int n = L.H * L.W;
        std::vector<double> a(size_t(n) * n, 0.0);
        for (int I = 0; I < L.H; ++I)
            for (int J = 0; J < L.W; ++J) {
                const int r = I * L.W + J;
                a[size_t(r) * n + r] = double(L.aP[r]);
                if (J + 1 < L.W) { a[size_t(r) * n + r + 1] = -double(L.cE[r]); a[size_t(r + 1) * n + r] = -double(L.cE[r]); }
                if (I + 1 < L.H) { a[size_t(r) * n + r + L.W] = -double(L.cN[r]); a[size_t(r + L.W) * n + r] = -double(L.cN[r]); }
            }
        gaussJordan(a.data(), n);
        inv.swap(a);
    }

    // In-place inverse; row k is scaled by 1/pivot, then eliminated from every
    // other row (the device kernel performs the same per-element updates).
    static void gaussJordan(double* a, int n) {
        std::vector<double> f(n);
        for (int k = 0; k < n; ++k) {
            double piv = a[size_t(k) * n + k];
            if (!(piv > 0.0)) piv = 1.0;
            const double ip = 1.0 / piv;
            a[size_t(k) * n + k] = 1.0;
            for (int j = 0; j < n; ++j) a[size_t(k) * n + j] *= ip;
            for (int i = 0; i < n; ++i) f[i] = (i == k) ? 0.0 : a[size_t(i) * n + k];
            for (int i = 0; i < n; ++i) {
                if (i == k) continue;
                a[size_t(i) * n + k] = 0.0;
                for (int j = 0; j < n; ++j) a[size_t(i) * n + j] -= f[i] * a[size_t(k) * n + j];
            }
        }
    }

    // Level 0 from the assembled CSR (DOF order is row-major over (I, J)).
    void build(const CSR<T>& A, int H, int W) {
        const size_t n = size_t(H) * W;
        std::vector<T> aP(n, T(0)), cE(n, T(0)), cN(n, T(0));
        for (int r = 0; r < A.n; ++r) {
            aP[r] = A.values[A.diagOff[r]];
            if (A.eOff[r] >= 0) cE[r] = -A.values[A.eOff[r]];
            if (A.nOff[r] >= 0) cN[r] = -A.values[A.nOff[r]];
        }
        buildGrid(H, W, aP.data(), cE.data(), cN.data());
    }

    void buildGrid(int H, int W, const T* aP, const T* cE, const T* cN) {
        lv.assign(1, Level());
        Level& L = lv[0];
        L.H = H;
        L.W = W;
        const size_t n = size_t(H) * W;
        L.aP.assign(aP, aP + n); L.cE.assign(cE, cE + n); L.cN.assign(cN, cN + n);
        L.b.assign(n, T(0)); L.x.assign(n, T(0)); L.z.assign(n, T(0));
        while ((lv.size() == 1 || size_t(lv.back().H) * lv.back().W > size_t(kCoarseMax)) &&
               int(lv.size()) < kMaxLevels) {
            Level c;
            coarsen(lv.back(), c);
            lv.push_back(std::move(c));
        }
        factor();
    }

    // x <- x + omega (b - A x) / aP over every unknown (Jacobi: reads old x)
    void sweep(Level& L, const std::vector<T>& xin, std::vector<T>& xout) {
        for (int I = 0; I < L.H; ++I)
            for (int J = 0; J < L.W; ++J) {
                const size_t k = L.at(I, J);
                xout[k] = xin[k] + omega * ((L.b[k] - ax(L, xin.data(), I, J)) / L.aP[k]);
            }
    }

    void vcycle(int l) {
        Level& L = lv[l];
        const size_t n = size_t(L.H) * L.W;
        if (l + 1 == int(lv.size())) {
            for (size_t r = 0; r < n; ++r) {
                double s = 0.0;
                for (size_t k = 0; k < n; ++k) s += inv[r * n + k] * double(L.b[k]);
                L.z[r] = T(s);
            }
            return;
        }
        std::vector<T> x1(n), x2(n);
        for (size_t k = 0; k < n; ++k) x1[k] = omega * (L.b[k] / L.aP[k]);
        sweep(L, x1, x2);
        Level& C = lv[l + 1];
        for (int I = 0; I < C.H; ++I)
            for (int J = 0; J < C.W; ++J) {
                T acc = T(0);
                for (int q = 0; q < 4; ++q) {
                    const int fi = 2 * I + (q >> 1), fj = 2 * J + (q & 1);
                    if (fi < L.H && fj < L.W) acc += L.b[L.at(fi, fj)] - ax(L, x2.data(), fi, fj);
                }
                C.b[C.at(I, J)] = acc;
            }
        vcycle(l + 1);
        for (int I = 0; I < L.H; ++I)
            for (int J = 0; J < L.W; ++J) x1[L.at(I, J)] = x2[L.at(I, J)] + scale * C.z[C.at(I / 2, J / 2)];
        sweep(L, x1, x2);
        sweep(L, x2, L.z);
    }

    void apply(const T* r, T* z) {
        Level& L = lv[0];
        std::copy(r, r + size_t(L.H) * L.W, L.b.begin());
        vcycle(0);
        std::copy(L.z.begin(), L.z.end(), z);
    }
};

// CG preconditioned by one aggregation-multigrid V-cycle; iteration, stop
// tests and return value as pcg() above.
template <typename T>
int pcgMG(const CSR<T>& A, AggMG<T>& mg, const T* b, T* x, T tol, int maxIter) {
    const int n = A.n;
    if (n == 0) return 0;
    std::vector<T> r(n), z(n), p(n), Ap(n);
    spmv(A, x, Ap.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    mg.apply(r.data(), z.data());
    p = z;
    T rz = blockedSum(n, r.data(), z.data());
    int iter = 0;
    for (; iter < maxIter; ++iter) {
        spmv(A, p.data(), Ap.data());
        const T pAp = blockedSum(n, p.data(), Ap.data());
        if (std::abs(pAp) < T(1e-20)) break;
        const T alpha = rz / pAp;
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        const T err = blockedSum(n, r.data(), r.data());
        if (std::sqrt(err) < tol) break;
        mg.apply(r.data(), z.data());
        const T rzPrev = rz;
        rz = blockedSum(n, r.data(), z.data());
        const T beta = rz / rzPrev;
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    return iter;
}

template <typename T>
struct PlaneT {
    T flowArea = 0, massFlux = 0, avgStatic = 0, avgDynamic = 0, avgTotal = 0;
    bool valid = false;
};

template <typename T>
struct Simple {
    cp_params prm{};
    int M = 0, N = 0;
    // parameters in working precision
    T hx, hy, rho, eta, targetVel, uvAlpha, pAlpha, Ht, sinkMult, alphaMax, q;
    T timeStep, minPseudoSpeed, pTol, momentumTol, epsilon;
    // run-time scalars (SIMPLE.h:260-368)
    T inletVelocity = 0, pseudoCFL = T(0.1), serResidPrev = 0, lsCurrentAlpha = 1;
    int lsBacktrackCount = 0;
    T residU = 1, residV = 1, residMass = 1, error = 1;
    int iter = 0;
    int lastItU = 0, lastItV = 0, lastItP = 0;
    // fields
    std::vector<T> u, v, p, uS, vS, pS, uOld, vOld, dE, dN, cellType, gamma, alpha;
    // persistent systems
    CSR<T> cU, cV, cP;
    AggMG<T> mg;
    std::vector<int> linU, linV, linP;

    // index helpers (reference shapes, row-major)
    T& U(std::vector<T>& a, int i, int j) { return a[size_t(i) * N + j]; }
    T& V(std::vector<T>& a, int i, int j) { return a[size_t(i) * (N + 1) + j]; }
    T& P(std::vector<T>& a, int i, int j) { return a[size_t(i) * (N + 1) + j]; }
    T& C(std::vector<T>& a, int i, int j) { return a[size_t(i) * N + j]; }
    T U(const std::vector<T>& a, int i, int j) const { return a[size_t(i) * N + j]; }
    T V(const std::vector<T>& a, int i, int j) const { return a[size_t(i) * (N + 1) + j]; }
    T P(const std::vector<T>& a, int i, int j) const { return a[size_t(i) * (N + 1) + j]; }
    T C(const std::vector<T>& a, int i, int j) const { return a[size_t(i) * N + j]; }

    static T clampv(T x, T lo, T hi) { return std::max(lo, std::min(hi, x)); }

    // SIMPLE.cpp:31-90 (parameter-derived values), 178-201 (fields),
    // 556-630 (gamma, Brinkman alpha). cellType is M x N row-major.
    void init(const cp_params& in, const double* ct) {
        prm = in;
        M = in.M;
        N = in.N;
        hx = T(in.hx); hy = T(in.hy); rho = T(in.rho); eta = T(in.mu);
        targetVel = T(in.U_in); uvAlpha = T(in.uvAlpha); pAlpha = T(in.pAlpha);
        Ht = T(in.Ht); sinkMult = T(in.sinkMultiplier); q = T(in.brinkmanQ);
        minPseudoSpeed = T(in.minPseudoSpeed);
        pTol = T(in.pTol); momentumTol = T(in.momentumTol); epsilon = T(in.epsilon);
        if (in.enableTwoPointFiveD && Ht <= T(0)) Ht = hy;  // SIMPLE.cpp:55-60
        if (in.alphaMax > 0) {
            alphaMax = T(in.alphaMax);
        } else {
            const T kmin = std::max(T(in.KMin), T(1e-20));  // SIMPLE.cpp:65-66
            alphaMax = eta / kmin;
        }
        const T domainLength = T(N) * hx;  // SIMPLE.cpp:68-70
        const T velRef = std::max(targetVel, T(1e-6));
        timeStep = T(in.timeStepMultiplier) * domainLength / velRef;
        pseudoCFL = in.enableCflRamp ? T(in.pseudoCFLInitial) : T(0.1);  // SIMPLE.cpp:79-82, SIMPLE.h:263
        u.assign(size_t(M + 1) * N, T(0));
        v.assign(size_t(M) * (N + 1), T(0));
        p.assign(size_t(M + 1) * (N + 1), T(0));
        uS = u; vS = v; pS = p; uOld = u; vOld = v; dE = u; dN = v;
        cellType.assign(size_t(M) * N, T(0));
        gamma.assign(size_t(M) * N, T(1));
        alpha.assign(size_t(M) * N, T(0));
        for (size_t k = 0; k < cellType.size(); ++k) cellType[k] = T(ct[k]);
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < N; ++j) C(gamma, i, j) = T(1) - C(cellType, i, j);  // SIMPLE.cpp:569
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < N; ++j) {
                const T g = C(gamma, i, j);
                C(alpha, i, j) = alphaMax * (T(1) - g) / (T(1) + q * g + T(1e-10));  // SIMPLE.cpp:608
            }
        if (M >= 2) {  // SIMPLE.cpp:616-625
            for (int j = 0; j < N; ++j) {
                C(alpha, 0, j) = alphaMax;
                C(alpha, M - 2, j) = alphaMax;
                C(gamma, 0, j) = T(0);
                C(gamma, M - 2, j) = T(0);
            }
        }
        inletVelocity = targetVel;
        buildSystems();
    }

    // alphaAtU / alphaAtV (SIMPLE.h:546-590).
    T alphaAtU(int i, int j) const {
        if (i < 1 || i >= M || j < 1 || j >= N - 1) return T(0);
        return T(0.5) * (C(alpha, i - 1, j - 1) + C(alpha, i - 1, j));
    }
    T alphaAtV(int i, int j) const {
        if (i < 1 || i >= M - 1 || j < 1 || j >= N) return T(0);
        return T(0.5) * (C(alpha, i - 1, j - 1) + C(alpha, i, j - 1));
    }

    // DOF maps and CSR patterns: momentum_solver.cpp:52-159 (U: i in
    // [1,M-1], j in [1,N-2]; V: i in [1,M-2], j in [1,N-1]) and
    // pressure_solver.cpp:438-458, 641-690 (p': i in [1,M-1], j in [1,N-2]).
    void buildOne(CSR<T>& c, std::vector<int>& lin, int rows, int cols, int i0, int i1, int j0, int j1) {
        lin.assign(size_t(rows) * cols, -1);
        c = CSR<T>();
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j) {
                lin[size_t(i) * cols + j] = c.n++;
                c.ri.push_back(i);
                c.rj.push_back(j);
            }
        c.row_ptr.assign(c.n + 1, 0);
        c.diagOff.assign(c.n, -1); c.eOff.assign(c.n, -1); c.wOff.assign(c.n, -1);
        c.nOff.assign(c.n, -1); c.sOff.assign(c.n, -1);
        auto at = [&](int i, int j) -> int {
            if (i < 0 || i >= rows || j < 0 || j >= cols) return -1;
            return lin[size_t(i) * cols + j];
        };
        int nnz = 0;
        for (int r = 0; r < c.n; ++r) {
            const int i = c.ri[r], j = c.rj[r];
            c.row_ptr[r] = nnz;
            // ascending column order: S, W, P, E, N
            const int s = at(i - 1, j), w = at(i, j - 1), e = at(i, j + 1), nn = at(i + 1, j);
            if (s >= 0) { c.sOff[r] = nnz; c.col_idx.push_back(s); nnz++; }
            if (w >= 0) { c.wOff[r] = nnz; c.col_idx.push_back(w); nnz++; }
            c.diagOff[r] = nnz; c.col_idx.push_back(r); nnz++;
            if (e >= 0) { c.eOff[r] = nnz; c.col_idx.push_back(e); nnz++; }
            if (nn >= 0) { c.nOff[r] = nnz; c.col_idx.push_back(nn); nnz++; }
        }
        c.row_ptr[c.n] = nnz;
        c.values.assign(nnz, T(0));
        c.rhs.assign(c.n, T(0));
        c.sol.assign(c.n, T(0));
    }
    void buildSystems() {
        buildOne(cU, linU, M + 1, N, 1, M - 1, 1, N - 2);
        buildOne(cV, linV, M, N + 1, 1, M - 2, 1, N - 1);
        buildOne(cP, linP, M + 1, N + 1, 1, M - 1, 1, N - 2);
    }

    // computeSOUCorrectionU / V (convection.cpp:53-186), index guards verbatim.
    T souU(int i, int j, T Fe, T Fw, T Fn, T Fs) const {
        T S = T(0);
        auto f = [&](int a, int b) { return U(u, a, b); };
        if (Fe >= T(0)) { if (j >= 2) S -= Fe * ((T(1.5) * f(i, j) - T(0.5) * f(i, j - 1)) - f(i, j)); }
        else { if (j + 2 < N) S -= Fe * ((T(1.5) * f(i, j + 1) - T(0.5) * f(i, j + 2)) - f(i, j + 1)); }
        if (Fw >= T(0)) { if (j >= 2) S += Fw * ((T(1.5) * f(i, j - 1) - T(0.5) * f(i, j - 2)) - f(i, j - 1)); }
        else { if (j + 1 < N) S += Fw * ((T(1.5) * f(i, j) - T(0.5) * f(i, j + 1)) - f(i, j)); }
        if (Fn >= T(0)) { if (i >= 1) S -= Fn * ((T(1.5) * f(i, j) - T(0.5) * f(i - 1, j)) - f(i, j)); }
        else { if (i + 2 <= M) S -= Fn * ((T(1.5) * f(i + 1, j) - T(0.5) * f(i + 2, j)) - f(i + 1, j)); }
        if (Fs >= T(0)) { if (i >= 2) S += Fs * ((T(1.5) * f(i - 1, j) - T(0.5) * f(i - 2, j)) - f(i - 1, j)); }
        else { if (i + 1 <= M) S += Fs * ((T(1.5) * f(i, j) - T(0.5) * f(i + 1, j)) - f(i, j)); }
        return S;
    }
    T souV(int i, int j, T Fe, T Fw, T Fn, T Fs) const {
        T S = T(0);
        auto f = [&](int a, int b) { return V(v, a, b); };
        if (Fe >= T(0)) { if (j >= 2) S -= Fe * ((T(1.5) * f(i, j) - T(0.5) * f(i, j - 1)) - f(i, j)); }
        else { if (j + 2 <= N) S -= Fe * ((T(1.5) * f(i, j + 1) - T(0.5) * f(i, j + 2)) - f(i, j + 1)); }
        if (Fw >= T(0)) { if (j >= 2) S += Fw * ((T(1.5) * f(i, j - 1) - T(0.5) * f(i, j - 2)) - f(i, j - 1)); }
        else { if (j + 1 <= N) S += Fw * ((T(1.5) * f(i, j) - T(0.5) * f(i, j + 1)) - f(i, j)); }
        if (Fn >= T(0)) { if (i >= 1) S -= Fn * ((T(1.5) * f(i, j) - T(0.5) * f(i - 1, j)) - f(i, j)); }
        else { if (i + 2 < M) S -= Fn * ((T(1.5) * f(i + 1, j) - T(0.5) * f(i + 2, j)) - f(i + 1, j)); }
        if (Fs >= T(0)) { if (i >= 2) S += Fs * ((T(1.5) * f(i - 1, j) - T(0.5) * f(i - 2, j)) - f(i - 1, j)); }
        else { if (i + 1 < M) S += Fs * ((T(1.5) * f(i, j) - T(0.5) * f(i + 1, j)) - f(i, j)); }
        return S;
    }

    // Momentum assembly into the persistent CSR (momentum_solver.cpp:185-298).
    template <bool IsU>
    void assembleMomentum() {
        CSR<T>& c = IsU ? cU : cV;
        const T vol = hx * hy;
        const T maxVel = T(3) * std::max(std::abs(inletVelocity), T(0.1));
        const T sinkCoeff = (prm.enableTwoPointFiveD && Ht > T(0))
                                ? sinkMult * (T(5) * eta / (T(2) * Ht * Ht)) : T(0);
        const T sinkDiag = sinkCoeff * vol;
        const T hChar = std::min(hx, hy);
        const T De = eta * hy / hx, Dn = eta * hx / hy;
        const T convScale = prm.enableConvectionScaling ? (T(6) / T(7)) : T(1);
#pragma omp parallel for schedule(static)
        for (int row = 0; row < c.n; ++row) {
            const int i = c.ri[row], j = c.rj[row];
            const T alphaLocal = IsU ? alphaAtU(i, j) : alphaAtV(i, j);
            T ue, uw, vn, vs;
            if (IsU) {
                ue = clampv(T(0.5) * (U(u, i, j) + U(u, i, j + 1)), -maxVel, maxVel);
                uw = clampv(T(0.5) * (U(u, i, j - 1) + U(u, i, j)), -maxVel, maxVel);
                vn = clampv(T(0.5) * (V(v, i, j) + V(v, i, j + 1)), -maxVel, maxVel);
                vs = clampv(T(0.5) * (V(v, i - 1, j) + V(v, i - 1, j + 1)), -maxVel, maxVel);
            } else {
                ue = clampv(T(0.5) * (U(u, i, j) + U(u, i + 1, j)), -maxVel, maxVel);
                uw = clampv(T(0.5) * (U(u, i, j - 1) + U(u, i + 1, j - 1)), -maxVel, maxVel);
                vn = clampv(T(0.5) * (V(v, i, j) + V(v, i + 1, j)), -maxVel, maxVel);
                vs = clampv(T(0.5) * (V(v, i - 1, j) + V(v, i, j)), -maxVel, maxVel);
            }
            const T Fe = convScale * rho * hy * ue, Fw = convScale * rho * hy * uw;
            const T Fn = convScale * rho * hx * vn, Fs = convScale * rho * hx * vs;
            const T aE = De + std::max(T(0), -Fe);
            const T aW = De + std::max(T(0), Fw);
            const T aN = Dn + std::max(T(0), -Fn);
            const T aS = Dn + std::max(T(0), Fs);
            const T sumA = aE + aW + aN + aS;
            T transCoeff = T(0);
            if (prm.enablePseudoTime) {
                T dtLocal = std::max(timeStep, T(1e-20));
                if (prm.useLocalPseudoTime) {
                    const T speed = IsU ? std::max(std::abs(U(u, i, j)), minPseudoSpeed)
                                        : std::max(std::abs(V(v, i, j)), minPseudoSpeed);
                    dtLocal = std::min(dtLocal, pseudoCFL * hChar / speed);
                    dtLocal = std::max(dtLocal, T(1e-20));
                }
                transCoeff = rho * vol / dtLocal;
            }
            const T brinkmanDrag = alphaLocal * vol;
            T aP0 = sumA + (Fe - Fw + Fn - Fs) + transCoeff + sinkDiag + brinkmanDrag;
            aP0 = std::max(aP0, T(1e-10));
            const T aP = aP0 / uvAlpha;
            T Sp, velOld, velCur;
            if (IsU) {
                Sp = (P(p, i, j) - P(p, i, j + 1)) * hy;
                velOld = U(uOld, i, j); velCur = U(u, i, j);
            } else {
                Sp = (P(p, i, j) - P(p, i + 1, j)) * hx;
                velOld = V(vOld, i, j); velCur = V(v, i, j);
            }
            Sp += transCoeff * velOld;
            Sp += (T(1) - uvAlpha) / uvAlpha * aP0 * velCur;
            if (prm.convectionScheme == 1) Sp += IsU ? souU(i, j, Fe, Fw, Fn, Fs) : souV(i, j, Fe, Fw, Fn, Fs);
            if (IsU) U(dE, i, j) = hy / aP0;
            else V(dN, i, j) = hx / aP0;
            auto upd = [&](int off, T coeff, T nbVal) {
                if (off >= 0) c.values[off] = -coeff;
                else Sp += coeff * nbVal;
            };
            c.values[c.diagOff[row]] = aP;
            if (IsU) {
                upd(c.eOff[row], aE, U(u, i, j + 1));
                upd(c.wOff[row], aW, U(u, i, j - 1));
                upd(c.nOff[row], aN, U(u, i + 1, j));
                upd(c.sOff[row], aS, U(u, i - 1, j));
                c.sol[row] = U(u, i, j);
            } else {
                upd(c.eOff[row], aE, V(v, i, j + 1));
                upd(c.wOff[row], aW, V(v, i, j - 1));
                upd(c.nOff[row], aN, V(v, i + 1, j));
                upd(c.sOff[row], aS, V(v, i - 1, j));
                c.sol[row] = V(v, i, j);
            }
            c.rhs[row] = Sp;
        }
    }

    // Assembly + Jacobi + clamp/scatter with max residual
    // (momentum_solver.cpp:164-339). Returns the sweep count.
    template <bool IsU>
    int solveMomentum(T& resid) {
        CSR<T>& c = IsU ? cU : cV;
        if (c.n == 0) { resid = T(0); return 0; }
        assembleMomentum<IsU>();
        const int it = jacobi<T>(c.n, c.row_ptr.data(), c.col_idx.data(), c.values.data(), c.rhs.data(),
                                 c.sol.data(), c.diagOff.data(), momentumTol, prm.maxMomentumIter);
        const T maxVel = T(3) * std::max(std::abs(inletVelocity), T(0.1));
        T r = T(0);
#pragma omp parallel for reduction(max : r) schedule(static)
        for (int row = 0; row < c.n; ++row) {
            const int i = c.ri[row], j = c.rj[row];
            const T nv = clampv(c.sol[row], -maxVel, maxVel);
            if (IsU) { r = std::max(r, std::abs(nv - U(u, i, j))); U(uS, i, j) = nv; }
            else { r = std::max(r, std::abs(nv - V(v, i, j))); V(vS, i, j) = nv; }
        }
        resid = r;
        return it;
    }

    // setVelocityBoundaryConditions (boundaries.cpp:48-106).
    void velocityBC(std::vector<T>& uu, std::vector<T>& vv) {
        const T Uin = inletVelocity;
        for (int j = 0; j < N; ++j) U(uu, 0, j) = T(0);
        for (int j = 0; j < N + 1; ++j) V(vv, 0, j) = T(0);
        for (int j = 0; j < N; ++j) U(uu, M, j) = T(0);
        for (int j = 0; j < N + 1; ++j) V(vv, M - 1, j) = T(0);
        for (int i = 1; i < M; ++i) {
            const int cr = i - 1;
            U(uu, i, 0) = (cr >= 0 && cr < M && C(gamma, cr, 0) > T(0.5)) ? Uin : T(0);
        }
        for (int i = 0; i < M; ++i) V(vv, i, 0) = T(0);
        for (int i = 0; i < M + 1; ++i) U(uu, i, N - 1) = U(uu, i, N - 2);
        for (int i = 0; i < M; ++i) V(vv, i, N) = V(vv, i, N - 1);
    }

    // setPressureBoundaryConditions (boundaries.cpp:111-134).
    void pressureBC(std::vector<T>& pp) {
        for (int j = 1; j < N; ++j) { P(pp, 0, j) = P(pp, 1, j); P(pp, M, j) = P(pp, M - 1, j); }
        for (int i = 1; i < M; ++i) P(pp, i, 0) = P(pp, i, 1);
        for (int i = 1; i < M; ++i) P(pp, i, N) = T(0);
    }

    // Pressure-correction coefficients of one p' DOF (pressure_solver.cpp:648-653).
    void presCoef(int i, int j, T& aE, T& aW, T& aN, T& aS, T& aP) const {
        aE = (j < N - 1) ? rho * U(dE, i, j) * hy : T(0);
        aW = (j > 1) ? rho * U(dE, i, j - 1) * hy : T(0);
        aN = (i < M - 1) ? rho * V(dN, i, j) * hx : T(0);
        aS = (i > 1) ? rho * V(dN, i - 1, j) * hx : T(0);
        aP = aE + aW + aN + aS;
        if (aP < T(1e-20)) aP = T(1);
    }

    // Mass imbalance of cell (i, j) (pressure_solver.cpp:476-480, 726-730).
    T massImbalance(int i, int j) const {
        const T mE = rho * U(uS, i, j) * hy, mW = rho * U(uS, i, j - 1) * hy;
        const T mN = rho * V(vS, i, j) * hx, mS = rho * V(vS, i - 1, j) * hx;
        return mW - mE + mS - mN;
    }

    // Mass residual + assembly + PCG + scatter (pressure_solver.cpp:462-806,
    // pressureSolverType 1). Returns the PCG iteration count.
    int solvePressure() {
        std::fill(pS.begin(), pS.end(), T(0));
        T rm = T(0);
        const int jEnd = prm.massResidualScope ? N - 1 : N;  // scope 1: p' columns only
#pragma omp parallel for reduction(max : rm) schedule(static)
        for (int i = 1; i < M; ++i)
            for (int j = 1; j < jEnd; ++j) rm = std::max(rm, std::abs(massImbalance(i, j)));
        residMass = rm;
        CSR<T>& c = cP;
        if (c.n == 0) return 0;
#pragma omp parallel for schedule(static)
        for (int r = 0; r < c.n; ++r) {
            const int i = c.ri[r], j = c.rj[r];
            T aE, aW, aN, aS, aP;
            presCoef(i, j, aE, aW, aN, aS, aP);
            c.values[c.diagOff[r]] = aP;
            if (c.eOff[r] >= 0) c.values[c.eOff[r]] = -aE;
            if (c.wOff[r] >= 0) c.values[c.wOff[r]] = -aW;
            if (c.nOff[r] >= 0) c.values[c.nOff[r]] = -aN;
            if (c.sOff[r] >= 0) c.values[c.sOff[r]] = -aS;
            c.rhs[r] = massImbalance(i, j);
            c.sol[r] = P(pS, i, j);
        }
        int it;
        if (prm.pressurePrecond == 1) {
            mg.omega = T(prm.mgOmega);
            mg.scale = T(prm.mgScale);
            mg.build(c, M - 1, N - 2);
            it = pcgMG<T>(c, mg, c.rhs.data(), c.sol.data(), pTol, prm.maxPressureIter);
        } else {
            it = pcg<T>(c, c.rhs.data(), c.sol.data(), pTol, prm.maxPressureIter);
        }
        for (int r = 0; r < c.n; ++r) P(pS, c.ri[r], c.rj[r]) = c.sol[r];
        for (int i = 1; i < M; ++i) P(pS, i, N - 1) = T(0);
        return it;
    }

    // p update, pressure BCs, velocity correction, velocity BCs
    // (iterations.cpp:456-500).
    void correct() {
        const T maxVel = T(3) * std::max(std::abs(inletVelocity), T(0.1));
        for (int i = 1; i < M; ++i)
            for (int j = 1; j < N; ++j) P(p, i, j) += pAlpha * P(pS, i, j);
        pressureBC(p);
#pragma omp parallel for schedule(static)
        for (int i = 1; i < M; ++i)
            for (int j = 1; j < N - 1; ++j) {
                T du = U(dE, i, j) * (P(pS, i, j) - P(pS, i, j + 1));
                du = clampv(du, -maxVel, maxVel);
                U(u, i, j) = clampv(U(uS, i, j) + du, -maxVel, maxVel);
            }
#pragma omp parallel for schedule(static)
        for (int i = 1; i < M - 1; ++i)
            for (int j = 1; j < N; ++j) {
                T dv = V(dN, i, j) * (P(pS, i, j) - P(pS, i + 1, j));
                dv = clampv(dv, -maxVel, maxVel);
                V(v, i, j) = clampv(V(vS, i, j) + dv, -maxVel, maxVel);
            }
        velocityBC(u, v);
    }

    // Explicit momentum sweep of one component (iterations.cpp:255-417,
    // momentumSolverType 0): every unknown updated once from the previous
    // iterate; returns max |new - old| over the unknowns.
    template <bool IsU>
    T explicitMomentum() {
        const T vol = hx * hy;
        const T maxVel = T(3) * std::max(std::abs(inletVelocity), T(0.1));
        const T sinkCoeff = (prm.enableTwoPointFiveD && Ht > T(0))
                                ? sinkMult * (T(5) * eta / (T(2) * Ht * Ht)) : T(0);
        const T sinkDiag = sinkCoeff * vol;
        const T hChar = std::min(hx, hy);
        const T De = eta * hy / hx, Dn = eta * hx / hy;
        const T convScale = prm.enableConvectionScaling ? (T(6) / T(7)) : T(1);
        const int iEnd = IsU ? M : M - 1, jEnd = IsU ? N - 1 : N;
        T r = T(0);
#pragma omp parallel for reduction(max : r) schedule(static)
        for (int i = 1; i < iEnd; ++i)
            for (int j = 1; j < jEnd; ++j) {
                const T alphaLocal = IsU ? alphaAtU(i, j) : alphaAtV(i, j);
                T ue, uw, vn, vs, cur, east, west, north, south;
                if (IsU) {
                    ue = clampv(T(0.5) * (U(u, i, j) + U(u, i, j + 1)), -maxVel, maxVel);
                    uw = clampv(T(0.5) * (U(u, i, j - 1) + U(u, i, j)), -maxVel, maxVel);
                    vn = clampv(T(0.5) * (V(v, i, j) + V(v, i, j + 1)), -maxVel, maxVel);
                    vs = clampv(T(0.5) * (V(v, i - 1, j) + V(v, i - 1, j + 1)), -maxVel, maxVel);
                    cur = U(u, i, j);
                    east = U(u, i, j + 1); west = U(u, i, j - 1); north = U(u, i + 1, j); south = U(u, i - 1, j);
                } else {
                    ue = clampv(T(0.5) * (U(u, i, j) + U(u, i + 1, j)), -maxVel, maxVel);
                    uw = clampv(T(0.5) * (U(u, i, j - 1) + U(u, i + 1, j - 1)), -maxVel, maxVel);
                    vn = clampv(T(0.5) * (V(v, i, j) + V(v, i + 1, j)), -maxVel, maxVel);
                    vs = clampv(T(0.5) * (V(v, i - 1, j) + V(v, i, j)), -maxVel, maxVel);
                    cur = V(v, i, j);
                    east = V(v, i, j + 1); west = V(v, i, j - 1); north = V(v, i + 1, j); south = V(v, i - 1, j);
                }
                const T Fe = convScale * rho * hy * ue, Fw = convScale * rho * hy * uw;
                const T Fn = convScale * rho * hx * vn, Fs = convScale * rho * hx * vs;
                const T aE = De + std::max(T(0), -Fe);
                const T aW = De + std::max(T(0), Fw);
                const T aN = Dn + std::max(T(0), -Fn);
                const T aS = Dn + std::max(T(0), Fs);
                const T Sdc = prm.convectionScheme == 1 ? (IsU ? souU(i, j, Fe, Fw, Fn, Fs) : souV(i, j, Fe, Fw, Fn, Fs))
                                                        : T(0);
                const T sumA = aE + aW + aN + aS;
                // computeLocalDt (iterations.cpp:210-216): no pseudo-time -> infinite step
                T transCoeff = T(0);
                if (prm.enablePseudoTime) {
                    T dtLocal = timeStep;
                    if (prm.useLocalPseudoTime) {
                        const T speed = std::max(std::abs(cur), minPseudoSpeed);
                        dtLocal = std::min(timeStep, pseudoCFL * hChar / speed);
                    }
                    transCoeff = rho * vol / dtLocal;
                }
                const T brinkmanDrag = alphaLocal * vol;
                T aP0 = sumA + (Fe - Fw + Fn - Fs) + transCoeff + sinkDiag + brinkmanDrag;
                aP0 = std::max(aP0, T(1e-10));
                const T dDenom = std::max(aP0, T(1e-12));
                T Sp;
                if (IsU) {
                    U(dE, i, j) = hy / dDenom;
                    Sp = (P(p, i, j) - P(p, i, j + 1)) * hy;
                } else {
                    V(dN, i, j) = hx / dDenom;
                    Sp = (P(p, i, j) - P(p, i + 1, j)) * hx;
                }
                Sp += transCoeff * cur;  // uOld == u here
                Sp += Sdc;
                const T aP = aP0 / uvAlpha;
                Sp += (T(1) - uvAlpha) / uvAlpha * aP0 * cur;
                T nv = (aE * east + aW * west + aN * north + aS * south + Sp) / aP;
                nv = clampv(nv, -maxVel, maxVel);
                r = std::max(r, std::abs(nv - cur));
                if (IsU) U(uS, i, j) = nv;
                else V(vS, i, j) = nv;
            }
        return r;
    }

    // Explicit momentum half of calculateStep (iterations.cpp:222-441).
    void momentumExplicit() {
        uOld = u;
        vOld = v;
        residU = explicitMomentum<true>();
        residV = explicitMomentum<false>();
        lastItU = lastItV = 0;
        velocityBC(uS, vS);
    }

    // Over-relaxation factor of the SOR backend (pressure_solver.cpp:810-816).
    static double sorOmegaFor(double w, int M, int N) {
        if (w > 0.0 && w < 2.0) return w;
        const double rs = std::cos(M_PI / double(std::max(M, N)));
        const double om = 2.0 / (1.0 + std::sqrt(1.0 - rs * rs));
        return std::max(1.0, std::min(1.95, om));
    }

    // Mass residual + red-black SOR on p' (pressure_solver.cpp:462-486,
    // 809-909, pressureSolverType 0). Returns the iteration count.
    int solvePressureSOR() {
        std::fill(pS.begin(), pS.end(), T(0));
        T rm = T(0);
        const int jEnd = prm.massResidualScope ? N - 1 : N;
        for (int i = 1; i < M; ++i)
            for (int j = 1; j < jEnd; ++j) rm = std::max(rm, std::abs(massImbalance(i, j)));
        residMass = rm;
        const T omega = T(sorOmegaFor(prm.sorOmega, M, N));
        int it = 0;
        for (int pIter = 0; pIter < prm.maxPressureIter; ++pIter) {
            T maxChange = T(0);
            it = pIter + 1;
            for (int color = 0; color < 2; ++color)
                for (int i = 1; i < M; ++i)
                    for (int j = 1; j < N; ++j) {
                        if ((i + j) % 2 != color) continue;
                        const T aE = (j < N - 1) ? rho * U(dE, i, j) * hy : T(0);
                        const T aW = (j > 1) ? rho * U(dE, i, j - 1) * hy : T(0);
                        const T aN = (i < M - 1) ? rho * V(dN, i, j) * hx : T(0);
                        const T aS = (i > 1) ? rho * V(dN, i - 1, j) * hx : T(0);
                        const T aP = aE + aW + aN + aS;
                        if (j == N - 1 || aP < T(1e-20)) {
                            P(pS, i, j) = T(0);
                            continue;
                        }
                        const T b = massImbalance(i, j);
                        const T pE = (j < N - 1) ? P(pS, i, j + 1) : T(0);
                        const T pW = (j > 1) ? P(pS, i, j - 1) : T(0);
                        const T pN = (i < M - 1) ? P(pS, i + 1, j) : T(0);
                        const T pSo = (i > 1) ? P(pS, i - 1, j) : T(0);
                        const T pGS = (aE * pE + aW * pW + aN * pN + aS * pSo + b) / aP;
                        const T pc = P(pS, i, j);
                        T pNew = pc + omega * (pGS - pc);
                        pNew = std::max(T(-10000), std::min(T(10000), pNew));
                        maxChange = std::max(maxChange, std::abs(pNew - pc));
                        P(pS, i, j) = pNew;
                    }
            if (maxChange < pTol) break;
        }
        return it;
    }

    // Momentum half of calculateStep (iterations.cpp:118-135).
    void momentum() {
        uOld = u;
        vOld = v;
        T ru, rv;
        lastItU = solveMomentum<true>(ru);
        lastItV = solveMomentum<false>(rv);
        residU = ru;
        residV = rv;
        velocityBC(uS, vS);
    }

    // calculateStep (iterations.cpp:82-526): momentum (implicit or explicit),
    // pressure correction (CG or SOR), correction.
    T calculateStep() {
        if (prm.momentumSolverType == 0) momentumExplicit();
        else momentum();
        lastItP = prm.pressureSolverType == 0 ? solvePressureSOR() : solvePressure();
        correct();
        return std::max({residU, residV, residMass});
    }

    // SER / line search / CFL ramp (time_control.cpp:43-165).
    void updateCflSER(T cur, int iteration) {
        if (!prm.enablePseudoTime || !prm.enableSER) return;
        if (iteration < prm.serMinIter) return;
        cur = std::max(cur, T(1e-15));
        if (serResidPrev <= T(0)) { serResidPrev = cur; return; }
        T ratio = serResidPrev / cur;
        ratio = std::max(ratio, T(prm.serMaxDecrease));
        ratio = std::min(ratio, T(prm.serMaxIncrease));
        T target = pseudoCFL * ratio;
        target = std::max(target, T(prm.serCFLMin));
        target = std::min(target, T(prm.serCFLMax));
        pseudoCFL = (T(1) - T(prm.serSmooth)) * pseudoCFL + T(prm.serSmooth) * target;
        if (!std::isfinite(pseudoCFL) || pseudoCFL <= T(0)) pseudoCFL = T(prm.serCFLMin);
        serResidPrev = cur;
    }
    void applyLineSearch(T cur) {
        if (!prm.enablePseudoTime || !prm.enableLineSearch) return;
        if (!prm.enableSER) return;
        if (serResidPrev <= T(0)) { lsCurrentAlpha = T(1); lsBacktrackCount = 0; return; }
        const T ratio = cur / serResidPrev;
        if (ratio > T(prm.lsResidIncreaseTol)) {
            lsBacktrackCount++;
            if (lsBacktrackCount <= prm.lsMaxTries) {
                lsCurrentAlpha *= T(prm.lsAlphaReduce);
                lsCurrentAlpha = std::max(lsCurrentAlpha, T(prm.lsAlphaMin));
                pseudoCFL *= T(prm.lsAlphaReduce);
                pseudoCFL = std::max(pseudoCFL, T(prm.serCFLMin));
            }
        } else {
            if (lsCurrentAlpha < T(1)) lsCurrentAlpha = std::min(lsCurrentAlpha * T(1.1), T(1));
            lsBacktrackCount = 0;
        }
    }
    void updateCflRamp(T cur) {
        if (!prm.enablePseudoTime || !prm.enableCflRamp) return;
        if (prm.enableSER) return;
        cur = std::max(cur, T(1e-12));
        if (cur > T(prm.cflRampStartRes)) {
            pseudoCFL = T(prm.pseudoCFLInitial);
        } else {
            const T ratio = T(prm.cflRampStartRes) / cur;
            const T raw = std::pow(ratio, T(prm.cflRampExponent));
            const T target = std::min(T(prm.pseudoCFLInitial) * raw, T(prm.pseudoCFLMax));
            pseudoCFL = (T(1) - T(prm.cflRampSmooth)) * pseudoCFL + T(prm.cflRampSmooth) * target;
            if (!std::isfinite(pseudoCFL) || pseudoCFL <= T(0)) pseudoCFL = T(prm.pseudoCFLInitial);
        }
    }

    // One pass of the outer loop body (SIMPLE.cpp:303-327), without I/O.
    T outerStep() {
        if (prm.enableInletRamp && prm.rampSteps > 0 && iter < prm.rampSteps)
            inletVelocity = (T(iter + 1) / T(prm.rampSteps)) * targetVel;
        else
            inletVelocity = targetVel;
        error = calculateStep();
        const T mon = prm.serUseMaxResid ? error : residMass;
        updateCflSER(mon, iter);
        applyLineSearch(mon);
        updateCflRamp(mon);
        iter++;
        return error;
    }

    // Pressure-drop convergence state (SIMPLE.cpp:266-294, 356-409).
    std::vector<T> dpHistory;
    int dpHistoryIdx = 0;
    bool dpHistoryFilled = false;
    bool dpConverged = false;
    std::vector<std::pair<int, T>> dpSlope;  // (iteration, core drop), oldest first

    // Least-squares slope of the core drop over the slope window
    // (SIMPLE.cpp:275-294).
    T dpSlopeValue() const {
        if (dpSlope.size() < 2) return T(0);
        const T x0 = T(dpSlope.front().first);
        T sx = 0, sy = 0, sxx = 0, sxy = 0;
        const T n = T(dpSlope.size());
        for (const auto& e : dpSlope) {
            const T x = T(e.first) - x0, y = e.second;
            sx += x;
            sy += y;
            sxx += x * x;
            sxy += x * y;
        }
        const T den = n * sxx - sx * sx;
        if (std::abs(den) < T(1e-12)) return T(0);
        return (n * sxy - sx * sy) / den;
    }

    // After outer iteration `it` (0-based, before the loop increments it):
    // core-plane total-pressure drop into the slope buffer and, from
    // iteration rampSteps on, into the moving window (SIMPLE.cpp:338-409).
    void monitorPressureDrop(int it) {
        const bool gate = prm.usePressureDropSlopeGate && prm.dpSlopeWindowIters > 0;
        if (!prm.usePressureDropConvergence && !gate) return;
        const T gmin = T(1) - T(1e-6);
        monitorDrop(it, plane(T(prm.xPlaneCoreInlet), gmin).avgTotal - plane(T(prm.xPlaneCoreOutlet), gmin).avgTotal);
    }

    // The window and slope rule applied to the core drop `dp` of iteration
    // `it` (SIMPLE.cpp:356-409).
    void monitorDrop(int it, T dp) {
        const bool gate = prm.usePressureDropSlopeGate && prm.dpSlopeWindowIters > 0;
        if (!prm.usePressureDropConvergence && !gate) return;
        if (gate) {
            dpSlope.emplace_back(it, dp);
            if (int(dpSlope.size()) > prm.dpSlopeWindowIters) dpSlope.erase(dpSlope.begin());
        }
        if (!prm.usePressureDropConvergence || it < prm.rampSteps) return;
        const int W = prm.dpConvergenceWindow;
        if (int(dpHistory.size()) != W) dpHistory.assign(W, T(0));
        dpHistory[dpHistoryIdx] = dp;
        dpHistoryIdx = (dpHistoryIdx + 1) % W;
        if (dpHistoryIdx == 0) dpHistoryFilled = true;
        if (!dpHistoryFilled) return;
        double sum = 0.0;
        for (int k = 0; k < W; ++k) sum += double(dpHistory[k]);
        const double mean = sum / double(W);
        double var = 0.0;
        for (int k = 0; k < W; ++k) {
            const double df = double(dpHistory[k]) - mean;
            var += df * df;
        }
        const double sd = std::sqrt(var / double(W));
        const double den = std::max(std::abs(mean), 1.0);
        const double sdPct = sd / den * 100.0;
        const T latest = dpHistory[(dpHistoryIdx - 1 + W) % W];
        const double latestPct = std::abs(double(latest) - mean) / den * 100.0;
        bool slopeOK = true;
        if (gate && int(dpSlope.size()) >= prm.dpSlopeWindowIters)
            slopeOK = std::abs(dpSlopeValue()) <= T(prm.dpSlopeMaxDelta) / T(prm.dpSlopeWindowIters);
        if (sdPct < prm.dpConvergencePercent && latestPct < prm.dpConvergencePercent && slopeOK) dpConverged = true;
    }

    // One pass of the runIterations loop body (SIMPLE.cpp:300-436, no I/O).
    T loopStep() {
        const int it = iter;
        outerStep();
        monitorPressureDrop(it);
        return error;
    }

    // runIterations loop condition (SIMPLE.cpp:299).
    int run(int maxIter) {
        int n = 0;
        while (iter < prm.maxIterations && error > epsilon && !dpConverged && n < maxIter) {
            loopStep();
            n++;
        }
        return n;
    }

    // samplePlaneAtX (postprocessing.cpp:53-231).
    PlaneT<T> plane(T xPhysical, T gammaMin) const {
        PlaneT<T> m;
        if (M <= 0 || N <= 0 || hx <= T(0) || hy <= T(0)) return m;
        const T xMin = T(0.5) * hx, xMax = (T(N) - T(0.5)) * hx;
        const T xc = std::max(xMin, std::min(xPhysical, xMax));
        const T jf = xc / hx - T(0.5);
        int cl = int(std::floor(jf));
        cl = std::max(0, std::min(cl, N - 2));
        const int cr = cl + 1;
        const T xl = (T(cl) + T(0.5)) * hx, xr = (T(cr) + T(0.5)) * hx;
        T t = (std::abs(xr - xl) < T(1e-12)) ? T(0) : (xc - xl) / (xr - xl);
        t = std::max(T(0), std::min(T(1), t));
        T sA = 0, dA = 0, tw = 0, area = 0, fs = 0;
        for (int i = 1; i < M; ++i) {
            const int row = i - 1;
            if (C(gamma, row, cl) < gammaMin || C(gamma, row, cr) < gammaMin) continue;
            const T fa = hy;
            area += fa;
            const int pl = std::max(0, std::min(cl + 1, N)), pr = std::max(0, std::min(cr + 1, N));
            const T sp = (T(1) - t) * P(p, i, pl) + t * P(p, i, pr);
            const int ul = std::max(0, std::min(cl, N - 1)), ur = std::max(0, std::min(cr, N - 1));
            const T un = (T(1) - t) * U(u, i, ul) + t * U(u, i, ur);
            const int vl = std::max(0, std::min(cl, N)), vr = std::max(0, std::min(cr, N));
            const int vt = std::max(0, std::min(i - 1, M - 1)), vb = std::max(0, std::min(i, M - 1));
            const T vT = (T(1) - t) * V(v, vt, vl) + t * V(v, vt, vr);
            const T vB = (T(1) - t) * V(v, vb, vl) + t * V(v, vb, vr);
            const T vtan = T(0.5) * (vT + vB);
            const T vm = std::sqrt(un * un + vtan * vtan);
            const T dyn = T(0.5) * rho * vm * vm;
            const T tot = sp + dyn;
            sA += sp * fa;
            dA += dyn * fa;
            const T fl = std::abs(rho * un * fa);
            if (fl > T(0)) { fs += fl; tw += tot * fl; }
        }
        m.flowArea = area;
        m.massFlux = fs;
        if (area > T(0)) { m.avgStatic = sA / area; m.avgDynamic = dA / area; m.valid = true; }
        if (fs > T(0)) m.avgTotal = tw / fs;
        else if (area > T(0)) m.avgTotal = m.avgStatic + m.avgDynamic;
        return m;
    }
};

}  // namespace oracle
