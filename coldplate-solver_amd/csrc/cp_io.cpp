// cp_io.cpp — the reference's file formats, host side: the parameter and
// geometry files it reads and the text / VTK files saveAll writes. Plain
// C++ over caller arrays (row-major, the reference's shapes); no device
// code, so the formats can be read and written without a GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/coldplate_hip.h"

namespace {

// Eigen's `out << float` with the stream defaults writes 6 significant
// digits (output.cpp:63); precision > 0 asks for more.
void putValue(std::ostream& out, double v, int precision) {
    if (precision > 0) out << std::setprecision(precision) << v;
    else out << float(v);
}

bool writeMatrix(const std::string& path, const double* a, int rows, int cols, int precision) {
    std::ofstream out(path);
    if (!out) return false;
    for (int i = 0; i < rows; ++i) {
        for (int j = 0; j < cols; ++j) {
            putValue(out, a[size_t(i) * cols + j], precision);
            if (j < cols - 1) out << '\t';
        }
        out << '\n';
    }
    return bool(out);
}

}  // namespace

extern "C" {

int cp_io_read_params(const char* path, cp_params* p, int* nInBuffer, int* nOutBuffer) {
    if (!path || !p) return CP_E_ARG;
    std::ifstream in(path);
    if (!in) return CP_E_ARG;
    int M = 0, N = 0, nIn = 0, nOut = 0;
    double hy = 0, hx = 0, u = 0;
    if (!(in >> M >> N >> hy >> hx >> u >> nIn >> nOut)) return CP_E_ARG;  // SIMPLE.cpp:39
    p->M = M;
    p->N = N;
    p->hy = hy;
    p->hx = hx;
    p->U_in = u;
    double ht = 0;
    if (in >> ht) p->Ht = ht;  // optional out-of-plane height (SIMPLE.cpp:44-47)
    // further columns (rho, nu) are ignored, as the reference does
    // (SIMPLE.cpp:49-54): fluid properties come from the parameters
    if (nInBuffer) *nInBuffer = nIn;
    if (nOutBuffer) *nOutBuffer = nOut;
    return CP_OK;
}

// Returns CP_OK, or 1 when the file is missing (every cell fluid, as
// SIMPLE.cpp:520-528). A file with fewer than M*N values leaves the rest
// fluid (0), as the reference's early return does (SIMPLE.cpp:531-537).
int cp_io_read_geometry(const char* path, int M, int N, double* cellType) {
    if (!path || !cellType || M < 1 || N < 1) return CP_E_ARG;
    for (size_t k = 0; k < size_t(M) * N; ++k) cellType[k] = 0.0;
    std::ifstream in(path);
    if (!in) return 1;
    for (size_t k = 0; k < size_t(M) * N; ++k) {
        double v;
        if (!(in >> v)) break;
        cellType[k] = v;
    }
    return CP_OK;
}

int cp_io_read_matrix(const char* path, int rows, int cols, double* a) {
    if (!path || !a) return CP_E_ARG;
    std::ifstream in(path);
    if (!in) return CP_E_ARG;
    for (size_t k = 0; k < size_t(rows) * cols; ++k)
        if (!(in >> a[k])) return CP_E_ARG;  // incomplete (SIMPLE.cpp:214-220)
    return CP_OK;
}

int cp_io_write_matrix(const char* path, const double* a, int rows, int cols, int precision) {
    if (!path || !a || rows < 0 || cols < 0) return CP_E_ARG;
    return writeMatrix(path, a, rows, cols, precision) ? CP_OK : CP_E_ARG;
}

// saveAll (output.cpp:73-266) into directory `dir`: u, v, p (staggered);
// u_full, v_full, pressure_full (cell-centred); u_thermal, v_thermal,
// pressure_thermal, p_gradient (columns nIn .. N-nOut-1); fluid_results.vtk.
int cp_io_save_all(const char* dir, int M, int N, double hx, double hy, int nIn, int nOut, const double* u,
                   const double* v, const double* p, const double* cellType, const double* gamma,
                   const double* alpha, int precision) {
    if (!dir || !u || !v || !p || !cellType || !gamma || !alpha || M < 2 || N < 2) return CP_E_ARG;
    const std::string d = std::string(dir) + "/";
    if (!writeMatrix(d + "u.txt", u, M + 1, N, precision)) return CP_E_ARG;
    if (!writeMatrix(d + "v.txt", v, M, N + 1, precision)) return CP_E_ARG;
    if (!writeMatrix(d + "p.txt", p, M + 1, N + 1, precision)) return CP_E_ARG;
    const size_t n = size_t(M) * N;
    std::vector<double> uc(n), vc(n), pc(n), gx(n), gy(n), gm(n);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) {
            const size_t k = size_t(i) * N + j;
            uc[k] = 0.5 * (u[size_t(i) * N + j] + u[size_t(i + 1) * N + j]);               // output.cpp:90
            vc[k] = 0.5 * (v[size_t(i) * (N + 1) + j] + v[size_t(i) * (N + 1) + j + 1]);  // output.cpp:91
            const int pr = std::max(1, std::min(i + 1, M - 1)), pcn = std::max(1, std::min(j + 1, N - 1));
            pc[k] = p[size_t(pr) * (N + 1) + pcn];  // output.cpp:95-97
        }
    if (!writeMatrix(d + "u_full.txt", uc.data(), M, N, precision)) return CP_E_ARG;
    if (!writeMatrix(d + "v_full.txt", vc.data(), M, N, precision)) return CP_E_ARG;
    if (!writeMatrix(d + "pressure_full.txt", pc.data(), M, N, precision)) return CP_E_ARG;
    auto P = [&](int i, int j) { return pc[size_t(i) * N + j]; };
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) {
            const size_t k = size_t(i) * N + j;
            // central differences inside, one-sided at the edges (output.cpp:126-147)
            gx[k] = (j > 0 && j < N - 1) ? (P(i, j + 1) - P(i, j - 1)) / (2.0 * hx)
                    : (j == 0)           ? (P(i, j + 1) - P(i, j)) / hx
                                         : (P(i, j) - P(i, j - 1)) / hx;
            gy[k] = (i > 0 && i < M - 1) ? (P(i + 1, j) - P(i - 1, j)) / (2.0 * hy)
                    : (i == 0)           ? (P(i + 1, j) - P(i, j)) / hy
                                         : (P(i, j) - P(i - 1, j)) / hy;
            gm[k] = std::sqrt(gx[k] * gx[k] + gy[k] * gy[k]);
        }
    const int nt = N - nIn - nOut;
    if (nt > 0) {  // output.cpp:155-181
        std::vector<double> ut(size_t(M) * nt), vt(ut.size()), pt(ut.size()), gt(ut.size());
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < nt; ++j) {
                const size_t s = size_t(i) * N + j + nIn, t = size_t(i) * nt + j;
                ut[t] = uc[s];
                vt[t] = vc[s];
                pt[t] = pc[s];
                gt[t] = gm[s];
            }
        if (!writeMatrix(d + "u_thermal.txt", ut.data(), M, nt, precision)) return CP_E_ARG;
        if (!writeMatrix(d + "v_thermal.txt", vt.data(), M, nt, precision)) return CP_E_ARG;
        if (!writeMatrix(d + "pressure_thermal.txt", pt.data(), M, nt, precision)) return CP_E_ARG;
        if (!writeMatrix(d + "p_gradient.txt", gt.data(), M, nt, precision)) return CP_E_ARG;
    }
    std::ofstream vtk(d + "fluid_results.vtk");  // output.cpp:186-262
    if (!vtk) return CP_E_ARG;
    vtk << "# vtk DataFile Version 3.0\nSIMPLE CFD Results\nASCII\nDATASET STRUCTURED_POINTS\n";
    vtk << "DIMENSIONS " << N << " " << M << " 1\nORIGIN 0 0 0\n";
    vtk << "SPACING " << float(hx) << " " << float(hy) << " 1\nPOINT_DATA " << n << "\n";
    auto scalars = [&](const char* name, const std::vector<double>& a) {
        vtk << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
        for (size_t k = 0; k < n; ++k) {
            putValue(vtk, a[k], precision);
            vtk << '\n';
        }
    };
    auto vectors = [&](const char* name, const std::vector<double>& x, const std::vector<double>& y) {
        vtk << "VECTORS " << name << " double\n";
        for (size_t k = 0; k < n; ++k) {
            putValue(vtk, x[k], precision);
            vtk << ' ';
            putValue(vtk, y[k], precision);
            vtk << " 0.0\n";
        }
    };
    scalars("pressure", pc);
    vectors("velocity", uc, vc);
    scalars("cellType", std::vector<double>(cellType, cellType + n));
    scalars("Density", std::vector<double>(gamma, gamma + n));
    scalars("Alpha", std::vector<double>(alpha, alpha + n));
    scalars("PressureGradient", gm);
    vectors("pressure_gradient", gx, gy);
    return vtk ? CP_OK : CP_E_ARG;
}

}  // extern "C"
