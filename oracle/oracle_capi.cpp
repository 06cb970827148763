// oracle_capi.cpp — C entry points of the CPU oracle (test infrastructure).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's CPU-baseline leg, as the checker. Never linked into the
// product library.
//
// Every function takes a `fp32` switch: 0 = FP64 oracle, 1 = the
// reference's all-float arithmetic (CPU baseline).
#include "simple_oracle.hpp"
#include "thermal_oracle.hpp"
#include "../include/coldplate_defaults.h"

#include <chrono>
#include <new>

using oracle::Simple;

namespace {

struct Handle {
    int fp32 = 0;
    Simple<double>* d = nullptr;
    Simple<float>* f = nullptr;
};

template <typename T>
void copyOut(const std::vector<T>& a, double* out) {
    for (size_t k = 0; k < a.size(); ++k) out[k] = double(a[k]);
}
template <typename T>
void copyIn(std::vector<T>& a, const double* in) {
    for (size_t k = 0; k < a.size(); ++k) a[k] = T(in[k]);
}

template <typename S>
void fillStats(const S& s, cp_step_stats* st) {
    if (!st) return;
    st->error = double(s.error);
    st->residU = double(s.residU);
    st->residV = double(s.residV);
    st->residMass = double(s.residMass);
    st->pseudoCFL = double(s.pseudoCFL);
    st->itU = s.lastItU;
    st->itV = s.lastItV;
    st->itP = s.lastItP;
    st->iter = s.iter;
    st->converged = (s.error <= s.epsilon || s.dpConverged) ? 1 : 0;
    st->health = 0;
}

template <typename T>
const std::vector<T>* fieldOf(const Simple<T>& s, int which) {
    switch (which) {
        case CP_F_U: return &s.u;
        case CP_F_V: return &s.v;
        case CP_F_P: return &s.p;
        case CP_F_USTAR: return &s.uS;
        case CP_F_VSTAR: return &s.vS;
        case CP_F_PSTAR: return &s.pS;
        case CP_F_DE: return &s.dE;
        case CP_F_DN: return &s.dN;
        case CP_F_GAMMA: return &s.gamma;
        case CP_F_ALPHA: return &s.alpha;
        default: return nullptr;
    }
}

template <typename T>
int getSystem(Simple<T>& s, int which, int* n, int* nnz, int* row_ptr, int* col_idx, double* values,
              double* rhs, double* sol, int* ri, int* rj) {
    const oracle::CSR<T>* c = which == 0 ? &s.cU : which == 1 ? &s.cV : which == 2 ? &s.cP : nullptr;
    if (!c) return -1;
    if (n) *n = c->n;
    if (nnz) *nnz = int(c->values.size());
    if (row_ptr) std::copy(c->row_ptr.begin(), c->row_ptr.end(), row_ptr);
    if (col_idx) std::copy(c->col_idx.begin(), c->col_idx.end(), col_idx);
    if (values) for (size_t k = 0; k < c->values.size(); ++k) values[k] = double(c->values[k]);
    if (rhs) for (int k = 0; k < c->n; ++k) rhs[k] = double(c->rhs[k]);
    if (sol) for (int k = 0; k < c->n; ++k) sol[k] = double(c->sol[k]);
    if (ri) std::copy(c->ri.begin(), c->ri.end(), ri);
    if (rj) std::copy(c->rj.begin(), c->rj.end(), rj);
    return 0;
}

template <typename T>
int pcgWrap(int n, const int* row_ptr, const int* col_idx, const T* values, const T* b, T* x, T tol,
                   int maxIter) {
    oracle::CSR<T> A;
    A.n = n;
    A.row_ptr.assign(row_ptr, row_ptr + n + 1);
    A.col_idx.assign(col_idx, col_idx + row_ptr[n]);
    A.values.assign(values, values + row_ptr[n]);
    return oracle::pcg<T>(A, b, x, tol, maxIter);
}
}  // namespace

#define DISPATCH(h, expr)            \
    do {                             \
        if ((h)->fp32) {             \
            auto& s = *(h)->f;       \
            expr;                    \
        } else {                     \
            auto& s = *(h)->d;       \
            expr;                    \
        }                            \
    } while (0)

extern "C" {

void or_params_default(cp_params* p) { cp_fill_reference_defaults(p); }

void* or_create(const cp_params* prm, const double* cellType, int fp32) {
    if (!prm || !cellType || prm->M < 4 || prm->N < 4) return nullptr;
    Handle* h = new (std::nothrow) Handle();
    if (!h) return nullptr;
    h->fp32 = fp32;
    if (fp32) { h->f = new Simple<float>(); h->f->init(*prm, cellType); }
    else { h->d = new Simple<double>(); h->d->init(*prm, cellType); }
    return h;
}

void or_destroy(void* hp) {
    Handle* h = static_cast<Handle*>(hp);
    if (!h) return;
    delete h->d;
    delete h->f;
    delete h;
}

int or_step(void* hp, cp_step_stats* st) {
    Handle* h = static_cast<Handle*>(hp);
    DISPATCH(h, { s.loopStep(); fillStats(s, st); });
    return 0;
}

// Runs up to max_steps outer iterations (loop condition of SIMPLE.cpp:299);
// returns the number of steps taken. wall_s receives the wall time.
int or_run(void* hp, int max_steps, cp_step_stats* st, double* wall_s) {
    Handle* h = static_cast<Handle*>(hp);
    int n = 0;
    const auto t0 = std::chrono::steady_clock::now();
    DISPATCH(h, { n = s.run(max_steps); fillStats(s, st); });
    const auto t1 = std::chrono::steady_clock::now();
    if (wall_s) *wall_s = std::chrono::duration<double>(t1 - t0).count();
    return n;
}

int or_momentum(void* hp, cp_step_stats* st) {
    Handle* h = static_cast<Handle*>(hp);
    DISPATCH(h, {
        if (s.prm.momentumSolverType == 0) s.momentumExplicit();
        else s.momentum();
        fillStats(s, st);
    });
    return 0;
}

int or_pressure(void* hp, cp_step_stats* st) {
    Handle* h = static_cast<Handle*>(hp);
    DISPATCH(h, {
        s.lastItP = s.prm.pressureSolverType == 0 ? s.solvePressureSOR() : s.solvePressure();
        fillStats(s, st);
    });
    return 0;
}

int or_correct(void* hp) {
    Handle* h = static_cast<Handle*>(hp);
    DISPATCH(h, { s.correct(); });
    return 0;
}

int or_get_field(void* hp, int which, double* out) {
    Handle* h = static_cast<Handle*>(hp);
    int rc = 0;
    DISPATCH(h, {
        auto* f = fieldOf(s, which);
        if (!f) rc = -1;
        else copyOut(*f, out);
    });
    return rc;
}

int or_set_fields(void* hp, const double* u, const double* v, const double* p) {
    Handle* h = static_cast<Handle*>(hp);
    DISPATCH(h, {
        if (u) { copyIn(s.u, u); s.uS = s.u; s.uOld = s.u; }
        if (v) { copyIn(s.v, v); s.vS = s.v; s.vOld = s.v; }
        if (p) copyIn(s.p, p);
    });
    return 0;
}

// Overwrites the current inlet velocity and pseudo-CFL (test hook).
int or_set_scalars(void* hp, double inletVelocity, double pseudoCFL) {
    Handle* h = static_cast<Handle*>(hp);
    DISPATCH(h, {
        s.inletVelocity = inletVelocity;
        s.pseudoCFL = pseudoCFL;
    });
    return 0;
}

int or_plane(void* hp, double x, double gammaMin, cp_plane_metrics* out) {
    Handle* h = static_cast<Handle*>(hp);
    DISPATCH(h, {
        using T = std::decay_t<decltype(s.hx)>;
        auto m = s.plane(T(x), T(gammaMin));
        out->flowArea = m.flowArea;
        out->massFlux = m.massFlux;
        out->avgStatic = m.avgStatic;
        out->avgDynamic = m.avgDynamic;
        out->avgTotal = m.avgTotal;
        out->valid = m.valid ? 1 : 0;
    });
    return 0;
}

// Exports the last assembled system: which 0 = U momentum, 1 = V momentum,
// 2 = pressure correction. Pass NULL arrays to query n / nnz first.
int or_get_system(void* hp, int which, int* n, int* nnz, int* row_ptr, int* col_idx, double* values,
                  double* rhs, double* sol, int* ri, int* rj) {
    Handle* h = static_cast<Handle*>(hp);
    int rc = 0;
    DISPATCH(h, { rc = getSystem(s, which, n, nnz, row_ptr, col_idx, values, rhs, sol, ri, rj); });
    return rc;
}

// The pseudo-CFL controller alone: for k = 0 .. n-1, the updates of
// SIMPLE.cpp:321-327 in that order (updateCflSER with iteration k,
// applyLineSearch, updateCflRamp; time_control.cpp:43-165) on monitored
// residual monitor[k], from the initial CFL of SIMPLE.cpp:79-82. Writes the
// CFL after each iteration and, when lsAlpha is non-null, the line search's
// step size.
int or_cfl_sequence(const cp_params* prm, int n, const double* monitor, double* cfl, double* lsAlpha) {
    if (!prm || n < 0 || (n > 0 && (!monitor || !cfl))) return -1;
    Simple<double> s;
    s.prm = *prm;
    s.pseudoCFL = prm->enableCflRamp ? prm->pseudoCFLInitial : 0.1;
    for (int k = 0; k < n; ++k) {
        s.updateCflSER(monitor[k], k);
        s.applyLineSearch(monitor[k]);
        s.updateCflRamp(monitor[k]);
        cfl[k] = s.pseudoCFL;
        if (lsAlpha) lsAlpha[k] = s.lsCurrentAlpha;
    }
    return 0;
}

// The pressure-drop window exit alone (SIMPLE.cpp:356-409) on given core
// drops dp[0 .. n-1] of iterations 0 .. n-1: returns the first iteration
// after which the run stops, or -1.
int or_dp_sequence(const cp_params* prm, int n, const double* dp) {
    if (!prm || n < 0 || (n > 0 && !dp)) return -2;
    Simple<double> s;
    s.prm = *prm;
    for (int k = 0; k < n; ++k) {
        s.monitorDrop(k, dp[k]);
        if (s.dpConverged) return k;
    }
    return -1;
}

// Thermal oracle: the reference's CSR system of the conjugate heat transfer
// solve (pass NULL arrays to query n and nnz first) and the statistics of a
// temperature field.
int or_thermal_assemble(const cp_thermal_params* p, const double* gamma, const double* u, const double* v, int* n,
                        int* nnz, int* row_ptr, int* col_idx, double* values, double* b) {
    if (!p || !gamma || !u || !v || p->Nx < 1 || p->Ny < 1) return -1;
    oracle::Thermal t;
    t.init(*p, gamma, u, v);
    const oracle::ThermalCSR A = t.assemble();
    if (n) *n = A.n;
    if (nnz) *nnz = int(A.values.size());
    if (row_ptr) std::copy(A.row_ptr.begin(), A.row_ptr.end(), row_ptr);
    if (col_idx) std::copy(A.col_idx.begin(), A.col_idx.end(), col_idx);
    if (values) std::copy(A.values.begin(), A.values.end(), values);
    if (b) std::copy(A.b.begin(), A.b.end(), b);
    return 0;
}

int or_thermal_stats(const cp_thermal_params* p, const double* gamma, const double* u, const double* v,
                     const double* T, cp_thermal_stats* st) {
    if (!p || !gamma || !u || !v || !T || !st) return -1;
    oracle::Thermal t;
    t.init(*p, gamma, u, v);
    t.stats(T, st);
    return 0;
}

void or_thermal_params_default(cp_thermal_params* p) { cp_fill_thermal_defaults(p); }

// Standalone restated solvers on caller CSR data (for the solver fixtures).
int or_jacobi_f32(int n, const int* row_ptr, const int* col_idx, const float* values, const float* b,
                  float* x, const int* diagOff, float tol, int maxIter) {
    return oracle::jacobi<float>(n, row_ptr, col_idx, values, b, x, diagOff, tol, maxIter);
}
int or_jacobi_f64(int n, const int* row_ptr, const int* col_idx, const double* values, const double* b,
                  double* x, const int* diagOff, double tol, int maxIter) {
    return oracle::jacobi<double>(n, row_ptr, col_idx, values, b, x, diagOff, tol, maxIter);
}

int or_pcg_f32(int n, const int* row_ptr, const int* col_idx, const float* values, const float* b, float* x,
               float tol, int maxIter) {
    return pcgWrap<float>(n, row_ptr, col_idx, values, b, x, tol, maxIter);
}
int or_pcg_f64(int n, const int* row_ptr, const int* col_idx, const double* values, const double* b,
               double* x, double tol, int maxIter) {
    return pcgWrap<double>(n, row_ptr, col_idx, values, b, x, tol, maxIter);
}

// Aggregation multigrid on a caller grid operator (H x W unknowns, row-major
// aP / cE / cN as AggMG): z = V-cycle(r). Returns the number of levels.
int or_mg_apply(int H, int W, const double* aP, const double* cE, const double* cN, double omega, double scale,
                const double* r, double* z) {
    oracle::AggMG<double> mg;
    mg.omega = omega;
    mg.scale = scale;
    mg.buildGrid(H, W, aP, cE, cN);
    mg.apply(r, z);
    return int(mg.lv.size());
}
int or_mg_apply_f32(int H, int W, const float* aP, const float* cE, const float* cN, float omega, float scale,
                    const float* r, float* z) {
    oracle::AggMG<float> mg;
    mg.omega = omega;
    mg.scale = scale;
    mg.buildGrid(H, W, aP, cE, cN);
    mg.apply(r, z);
    return int(mg.lv.size());
}
// Level `level` of the hierarchy: its shape (always) and, when the output
// arrays are non-null, its aP / cE / cN.
int or_mg_level(int H, int W, const double* aP, const double* cE, const double* cN, int level, int* Hl, int* Wl,
                double* aPl, double* cEl, double* cNl) {
    oracle::AggMG<double> mg;
    mg.buildGrid(H, W, aP, cE, cN);
    if (level < 0 || level >= int(mg.lv.size())) return -1;
    const auto& L = mg.lv[level];
    *Hl = L.H;
    *Wl = L.W;
    if (aPl) std::copy(L.aP.begin(), L.aP.end(), aPl);
    if (cEl) std::copy(L.cE.begin(), L.cE.end(), cEl);
    if (cNl) std::copy(L.cN.begin(), L.cN.end(), cNl);
    return int(mg.lv.size());
}

}  // extern "C"
