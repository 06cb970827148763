// thermal_oracle.hpp — CPU restatement of the reference's 3D conjugate heat
// transfer assembly and post-processing (ThermalSolver/thermal_solver.cpp).
//
// TEST INFRASTRUCTURE ONLY: compiled into oracle/lib/liboracle.so and used
// by tests/ as the checker of the device thermal solve. Never linked into
// the product library.
//
// What is restated, with the reference lines it follows:
//   * Params and the derived grid (thermal_solver.cpp:142-189);
//   * kval (RAMP / harmonic / linear conductivity, :260-290), region
//     (:240-244), pois (:311-315), harm (:235);
//   * the row-by-row COO assembly of ThermalSolver3D::solve (:421-518),
//     including the inlet Dirichlet term, the convective outlet terms, the
//     z-series conductances and the base heat flux;
//   * SparseMatrix::from_coo (:87-124): sort, then merge consecutive entries
//     of equal column;
//   * the temperature statistics and energy balance of solve (:556-604) and
//     the volume-weighted averages of compute_thermal_metrics
//     (thermal_output.cpp:135-205).
// The reference solves A T = b with AMGCL's IDR(8) + smoothed-aggregation
// AMG to a relative residual of 1e-6; AMGCL is not in the mount, and the
// solution of the linear system does not depend on the solver, so the tests
// solve the oracle's CSR system directly (scipy's sparse LU).
#pragma once

#include <algorithm>
#include <cmath>
#include <tuple>
#include <vector>

#include "../include/coldplate_hip.h"

namespace oracle {

struct ThermalCSR {
    int n = 0;
    std::vector<int> row_ptr, col_idx;
    std::vector<double> values, b;
};

struct Thermal {
    cp_thermal_params p{};
    std::vector<double> gamma, u, v, dz;

    int Nz() const { return p.nz_solid + p.nz_fluid; }
    int idx(int i, int j, int k) const { return k * p.Nx * p.Ny + j * p.Nx + i; }
    static double harm(double a, double b) { return 2 * a * b / (a + b + 1e-30); }

    void init(const cp_thermal_params& prm, const double* g, const double* uu, const double* vv) {
        p = prm;
        const size_t nc = size_t(p.Nx) * p.Ny;
        gamma.assign(g, g + nc);
        u.assign(uu, uu + nc);
        v.assign(vv, vv + nc);
        dz.assign(Nz(), 0.0);
        for (int k = 0; k < p.nz_solid; ++k) dz[k] = p.H_b / p.nz_solid;
        for (int k = 0; k < p.nz_fluid; ++k) dz[p.nz_solid + k] = p.H_t / p.nz_fluid;
    }

    int region(int i, int j, int k) const {
        if (k < p.nz_solid) return 0;
        return gamma[size_t(j) * p.Nx + i] > 0.01 ? 1 : 0;
    }

    double kval(int i, int j, int k) const {
        if (k < p.nz_solid) return p.k_s;
        double g = gamma[size_t(j) * p.Nx + i];
        g = std::max(0.0, std::min(1.0, g));
        switch (p.k_interp_mode) {
            case 1: return p.k_f + (p.k_s - p.k_f) * (1.0 - g) / (1.0 + p.qk * g);
            case 2: return p.k_s * p.k_f / ((1.0 - g) * p.k_f + g * p.k_s + 1e-30);
            default: return p.k_s * (1.0 - g) + p.k_f * g;
        }
    }

    double pois(int k) const {
        if (k < p.nz_solid) return 0;
        const double z = (k - p.nz_solid + 0.5) * (p.H_t / p.nz_fluid) / p.H_t;
        return 4.0 * z * (1 - z);
    }

    // The reference's COO rows, then from_coo.
    ThermalCSR assemble() const {
        const int Nx = p.Nx, Ny = p.Ny, Nz = this->Nz(), n = Nx * Ny * Nz;
        const double dx = p.dx, dy = p.dy, rhoCp = p.rho * p.Cp, Tin = p.T_inlet, Az = dx * dy;
        std::vector<std::tuple<int, int, double>> coo;
        ThermalCSR out;
        out.n = n;
        out.b.assign(n, 0.0);
        for (int k = 0; k < Nz; ++k) {
            const double dzk = dz[k], Ax = dy * dzk, Ay = dx * dzk;
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i) {
                    const int P = idx(i, j, k);
                    double diag = 0;
                    const double kP = kval(i, j, k), po = pois(k);
                    double g = (k < p.nz_solid) ? 0.0 : gamma[size_t(j) * Nx + i];
                    g = std::max(0.0, std::min(1.0, g));
                    const double uP = (g > 0.01) ? u[size_t(j) * Nx + i] * po : 0.0;
                    const double vP = (g > 0.01) ? v[size_t(j) * Nx + i] * po : 0.0;
                    const bool fl = region(i, j, k) == 1;
                    if (i > 0) {
                        const double D = harm(kP, kval(i - 1, j, k)) * Ax / dx;
                        diag += D;
                        coo.emplace_back(P, idx(i - 1, j, k), -D);
                        if (fl && region(i - 1, j, k) == 1) {
                            const double uW = u[size_t(j) * Nx + i - 1] * po, F = rhoCp * 0.5 * (uP + uW) * Ax;
                            diag += std::max(-F, 0.0);
                            coo.emplace_back(P, idx(i - 1, j, k), -std::max(F, 0.0));
                        }
                    } else if (fl) {
                        const double Dbc = 2.0 * kP * Ax / dx, Fbc = rhoCp * uP * Ax;
                        diag += Dbc;
                        out.b[P] += Dbc * Tin;
                        if (Fbc > 0) out.b[P] += Fbc * Tin;
                        else diag += -Fbc;
                    }
                    if (i < Nx - 1) {
                        const double D = harm(kP, kval(i + 1, j, k)) * Ax / dx;
                        diag += D;
                        coo.emplace_back(P, idx(i + 1, j, k), -D);
                        if (fl && region(i + 1, j, k) == 1) {
                            const double uE = u[size_t(j) * Nx + i + 1] * po, F = rhoCp * 0.5 * (uP + uE) * Ax;
                            diag += std::max(F, 0.0);
                            coo.emplace_back(P, idx(i + 1, j, k), -std::max(-F, 0.0));
                        }
                    } else if (fl && uP > 0) {
                        diag += rhoCp * uP * Ax;
                    }
                    if (j > 0) {
                        const double D = harm(kP, kval(i, j - 1, k)) * Ay / dy;
                        diag += D;
                        coo.emplace_back(P, idx(i, j - 1, k), -D);
                        if (fl && region(i, j - 1, k) == 1) {
                            const double vS = v[size_t(j - 1) * Nx + i] * po, F = rhoCp * 0.5 * (vP + vS) * Ay;
                            diag += std::max(-F, 0.0);
                            coo.emplace_back(P, idx(i, j - 1, k), -std::max(F, 0.0));
                        }
                    }
                    if (j < Ny - 1) {
                        const double D = harm(kP, kval(i, j + 1, k)) * Ay / dy;
                        diag += D;
                        coo.emplace_back(P, idx(i, j + 1, k), -D);
                        if (fl && region(i, j + 1, k) == 1) {
                            const double vN = v[size_t(j + 1) * Nx + i] * po, F = rhoCp * 0.5 * (vP + vN) * Ay;
                            diag += std::max(F, 0.0);
                            coo.emplace_back(P, idx(i, j + 1, k), -std::max(-F, 0.0));
                        }
                    } else if (fl && vP > 0) {
                        diag += rhoCp * vP * Ay;
                    }
                    if (k > 0) {
                        const double kN = kval(i, j, k - 1), dzP = dz[k], dzN = dz[k - 1];
                        const double kf = (dzP + dzN) / (dzP / kP + dzN / kN);
                        const double D = kf * Az / (0.5 * (dzP + dzN));
                        diag += D;
                        coo.emplace_back(P, idx(i, j, k - 1), -D);
                    } else {
                        out.b[P] += p.q0 * Az;
                    }
                    if (k < Nz - 1) {
                        const double kN = kval(i, j, k + 1), dzP = dz[k], dzN = dz[k + 1];
                        const double kf = (dzP + dzN) / (dzP / kP + dzN / kN);
                        const double D = kf * Az / (0.5 * (dzP + dzN));
                        diag += D;
                        coo.emplace_back(P, idx(i, j, k + 1), -D);
                    }
                    coo.emplace_back(P, P, diag + 1e-12);
                }
        }
        // from_coo: sort; consecutive entries of one row and column merge
        std::sort(coo.begin(), coo.end());
        out.row_ptr.assign(n + 1, 0);
        int last_r = -1, last_c = -1;
        for (const auto& [r, c, val] : coo) {
            if (r == last_r && c == last_c) {
                out.values.back() += val;
                continue;
            }
            while (last_r < r) out.row_ptr[++last_r] = int(out.col_idx.size());
            out.col_idx.push_back(c);
            out.values.push_back(val);
            last_c = c;
        }
        while (last_r < n) out.row_ptr[++last_r] = int(out.col_idx.size());
        return out;
    }

    // Statistics and energy balance of solve (thermal_solver.cpp:556-604) and
    // the averages of compute_thermal_metrics (thermal_output.cpp:135-205).
    void stats(const double* T, cp_thermal_stats* s) const {
        const int Nx = p.Nx, Ny = p.Ny, Nz = this->Nz();
        s->T_base_min = s->T_fin_min = s->T_fluid_min = 1e30;
        s->T_base_max = s->T_fin_max = s->T_fluid_max = -1e30;
        double tvB = 0, vB = 0, tvF = 0, vF = 0, tvS = 0, vS = 0, tvG = 0, vG = 0;
        for (int k = 0; k < Nz; ++k)
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i) {
                    const double t = T[idx(i, j, k)], V = p.dx * p.dy * dz[k];
                    tvG += t * V;
                    vG += V;
                    if (k < p.nz_solid) {
                        s->T_base_min = std::min(s->T_base_min, t);
                        s->T_base_max = std::max(s->T_base_max, t);
                        tvB += t * V;
                        vB += V;
                    } else if (region(i, j, k) == 1) {
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
        s->T_avg_base = vB > 0 ? tvB / vB : 0.0;
        s->T_avg_fluid = vF > 0 ? tvF / vF : 0.0;
        s->T_avg_solid = vS > 0 ? tvS / vS : 0.0;
        s->T_avg_global = vG > 0 ? tvG / vG : 0.0;
        double Qin = 0, mdot = 0, mdotT = 0;
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i) Qin += p.q0 * p.dx * p.dy;
        for (int k = p.nz_solid; k < Nz; ++k) {
            const double A = p.dy * dz[k], po = pois(k);
            for (int j = 0; j < Ny; ++j) {
                const int i = Nx - 1;
                if (region(i, j, k) == 1) {
                    const double uu = u[size_t(j) * Nx + i] * po;
                    if (uu > 0) {
                        const double m = p.rho * uu * A;
                        mdot += m;
                        mdotT += m * T[idx(i, j, k)];
                    }
                }
            }
        }
        s->T_out = mdot > 0 ? mdotT / mdot : p.T_inlet;
        s->Q_in = Qin;
        s->Q_out = mdot * p.Cp * (s->T_out - p.T_inlet);
        s->imbalance_pct = 100 * (Qin - s->Q_out) / (Qin + 1e-30);
    }
};

}  // namespace oracle
