// cp_geometry.cpp — seeded geometry generators for the solver's workloads.
//
// Output is a cellType field (M x N, row-major, 0 = fluid, 1 = solid,
// intermediate = Brinkman buffer), the format of the reference's
// geometry_fluid.txt (GeometryGeneratorV1.py:153, SIMPLE.cpp:517-537).
//
//  * straight: one channel on cell rows [r0, r1], solid elsewhere.
//  * network:  6-10 channels, 8-16 cells wide, routed from the inlet column
//    to the outlet column with Manhattan and 45-degree diagonal moves on a
//    4-cell lattice, with random splits; channels that meet merge. This is the
//    build's own generator after the idea of GeometryGeneratorV2.py
//    (default_rng routing, :98-154, splits/merges :258-371), not a copy: the
//    workloads here define new seeded families.
//  * gamma:    Gaussian random field (white noise smoothed by a separable
//    Gaussian of standard deviation `corr` cells), standardised, mapped to
//    [0,1] by the normal CDF and tanh-projected with (beta, eta); cellType is
//    1 - gamma.
// The random source is SplitMix64 (Steele, Lea, Flood 2014) with Box-Muller
// normals, so the fields are reproducible from the seed on any host.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/coldplate_hip.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    int range(int lo, int hi) { return lo + int(uniform() * double(hi - lo + 1)); }   // [lo,hi]
    double normal() {
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// Counter-based normal for the noise field: value depends only on (seed, idx).
double noiseAt(uint64_t seed, uint64_t idx) {
    Rng r(seed ^ (idx * 0xD1B54A32D192ED03ull));
    return r.normal();
}

void carve(std::vector<double>& ct, int M, int N, int ci, int cj, int w) {
    const int h = w / 2;
    for (int i = ci - h; i < ci - h + w; ++i) {
        if (i < 1 || i > M - 2) continue;
        for (int j = cj - h; j < cj - h + w; ++j) {
            if (j < 0 || j > N - 1) continue;
            ct[size_t(i) * N + j] = 0.0;
        }
    }
}

struct Path {
    int i, j, w;
};

}  // namespace

extern "C" {

int cp_geom_straight(int M, int N, int row0, int row1, double* ct) {
    if (!ct || M < 4 || N < 4 || row0 < 0 || row1 >= M || row0 > row1) return CP_E_ARG;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) ct[size_t(i) * N + j] = (i >= row0 && i <= row1) ? 0.0 : 1.0;
    return CP_OK;
}

int cp_geom_network(int M, int N, uint64_t seed, double* ct) {
    if (!ct || M < 16 || N < 16) return CP_E_ARG;
    std::vector<double> g(size_t(M) * N, 1.0);
    Rng rng(seed);
    const int step = 4;
    // widths scale down on small grids so that channels still fit
    const int wMax = std::max(2, std::min(16, (M - 4) / 6));
    const int wMin = std::max(2, std::min(8, wMax));
    const int nPaths = rng.range(6, 10);
    std::vector<Path> active;
    for (int k = 0; k < nPaths; ++k) {
        const int w = rng.range(wMin, wMax);
        const int lo = 2 + w / 2, hi = M - 3 - w / 2;
        const int i0 = lo + int(rng.uniform() * double(std::max(1, hi - lo)));
        active.push_back({std::min(std::max(i0, lo), hi), 0, w});
    }
    // advance every path column-wise until it reaches the outlet
    size_t guard = 0;
    while (!active.empty() && guard++ < size_t(16) * N) {
        std::vector<Path> next;
        for (Path pth : active) {
            const int lo = 2 + pth.w / 2, hi = M - 3 - pth.w / 2;
            carve(g, M, N, pth.i, pth.j, pth.w);
            const double r = rng.uniform();
            int di = 0, dj = step;
            if (r < 0.15) { di = -step; dj = step; }        // diagonal up
            else if (r < 0.30) { di = step; dj = step; }    // diagonal down
            else if (r < 0.40) { di = -step; dj = 0; }      // vertical (Manhattan) up
            else if (r < 0.50) { di = step; dj = 0; }       // vertical down
            const int ni = std::min(std::max(pth.i + di, lo), hi);
            const int nj = std::min(pth.j + dj, N - 1);
            // carve the segment cell by cell so diagonals stay connected
            const int n = std::max(std::abs(ni - pth.i), std::abs(nj - pth.j));
            for (int s = 1; s <= n; ++s) {
                const int ci = pth.i + (ni - pth.i) * s / std::max(n, 1);
                const int cj = pth.j + (nj - pth.j) * s / std::max(n, 1);
                carve(g, M, N, ci, cj, pth.w);
            }
            Path q{ni, nj, pth.w};
            if (nj >= N - 1) {
                carve(g, M, N, ni, N - 1, pth.w);
                continue;  // reached the outlet
            }
            // split with probability 0.05: a branch leaves at +-2 widths
            if (rng.uniform() < 0.05 && next.size() < 24) {
                const int off = (rng.uniform() < 0.5 ? -1 : 1) * 2 * pth.w;
                Path br{std::min(std::max(ni + off, lo), hi), nj, pth.w};
                const int bn = std::abs(br.i - ni);
                for (int s = 1; s <= bn; ++s) carve(g, M, N, ni + (br.i - ni) * s / std::max(bn, 1), nj, pth.w);
                next.push_back(br);
            }
            next.push_back(q);
        }
        active.swap(next);
    }
    // walls: top and bottom cell rows solid
    for (int j = 0; j < N; ++j) {
        g[j] = 1.0;
        g[size_t(M - 1) * N + j] = 1.0;
    }
    std::copy(g.begin(), g.end(), ct);
    return CP_OK;
}

int cp_geom_gamma(int M, int N, uint64_t seed, double corr, double beta, double eta, double* ct) {
    if (!ct || M < 4 || N < 4 || !(corr > 0) || !(beta > 0)) return CP_E_ARG;
    const int R = int(std::ceil(3.0 * corr));
    std::vector<double> ker(2 * R + 1);
    double ks = 0.0;
    for (int k = -R; k <= R; ++k) {
        ker[k + R] = std::exp(-0.5 * double(k) * double(k) / (corr * corr));
        ks += ker[k + R];
    }
    for (double& x : ker) x /= ks;
    std::vector<double> a(size_t(M) * N), b(size_t(M) * N);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) a[size_t(i) * N + j] = noiseAt(seed, uint64_t(i) * uint64_t(N) + uint64_t(j));
    // separable smoothing with periodic wrap
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) {
            double s = 0.0;
            for (int k = -R; k <= R; ++k) s += ker[k + R] * a[size_t(i) * N + ((j + k) % N + N) % N];
            b[size_t(i) * N + j] = s;
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) {
            double s = 0.0;
            for (int k = -R; k <= R; ++k) s += ker[k + R] * b[size_t((i + k) % M + M) % M * N + j];
            a[size_t(i) * N + j] = s;
        }
    double mean = 0.0, var = 0.0;
    for (double x : a) mean += x;
    mean /= double(a.size());
    for (double x : a) var += (x - mean) * (x - mean);
    const double sd = std::sqrt(var / double(a.size())) + 1e-300;
    const double t0 = std::tanh(beta * eta), t1 = std::tanh(beta * (1.0 - eta));
    for (size_t k = 0; k < a.size(); ++k) {
        const double z = (a[k] - mean) / sd;
        const double gm = 0.5 * std::erfc(-z / std::sqrt(2.0));  // normal CDF
        const double gp = (t0 + std::tanh(beta * (gm - eta))) / (t0 + t1);
        ct[k] = 1.0 - std::min(1.0, std::max(0.0, gp));
    }
    return CP_OK;
}

}  // extern "C"
