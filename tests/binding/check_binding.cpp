// Drives the reference-side binding against the SIMPLE stand-in: parameter
// mapping, Eigen <-> row-major conversion, and the error path of
// hipBackendInit. Prints one line per check; exits 1 on the first failure.
#include "SIMPLE.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "hip_backend.h"

template <class Mat>
std::vector<double> hipRowMajor(const Mat& a);
template <class Mat>
void hipFromRowMajor(const std::vector<double>& in, Mat& a);

static int fails = 0;
static void expect(bool ok, const char* what) {
    std::printf("%s %s\n", ok ? "ok" : "FAIL", what);
    if (!ok) ++fails;
}

int main(int argc, char** argv) {
    const bool expectDevice = argc > 1 && std::strcmp(argv[1], "--device") == 0;
    SIMPLE s;
    s.M = 6;
    s.N = 9;
    s.hx = 5e-5f;
    s.hy = 4e-5f;
    s.rho = 1000.0f;
    s.eta = 1e-3f;
    s.targetVel = 0.1f;
    s.convectionScheme = 1;
    s.enableTwoPointFiveD = true;
    s.Ht_channel = 2e-4f;
    s.enablePseudoTimeStepping = true;
    s.serCFLMax = 7.0f;
    s.maxPressureIter = 33;
    s.hipPressureAlgorithm = 0;
    s.momentumSolverType = 0;
    s.brinkmanAlphaMax = s.eta / s.brinkmanKMin;
    s.usePressureDropConvergence = false;
    const cp_params p = hipParamsOf(s);
    expect(p.M == 6 && p.N == 9 && p.hx == double(5e-5f) && p.hy == double(4e-5f), "grid");
    expect(p.rho == 1000.0 && p.mu == double(1e-3f) && p.U_in == double(0.1f), "fluid and inlet");
    expect(p.convectionScheme == 1 && p.enableTwoPointFiveD == 1 && p.Ht == double(2e-4f) &&
               p.sinkMultiplier == double(4.8f),
           "convection and 2.5D");
    expect(p.enablePseudoTime == 1 && p.serCFLMax == 7.0 && p.serMinIter == 10 && p.lsMaxTries == 4, "pseudo-time");
    expect(p.maxPressureIter == 33 && p.pTol == double(1e-4f) && p.maxMomentumIter == 200, "inner solvers");
    expect(p.pressureSolverType == 0 && p.momentumSolverType == 0 && p.precision == CP_PREC_FP64, "algorithms");
    expect(p.alphaMax == double(s.brinkmanAlphaMax) && p.usePressureDropConvergence == 0, "Brinkman and exit");

    s.u = Eigen::MatrixXf(s.M + 1, s.N);
    s.v = Eigen::MatrixXf(s.M, s.N + 1);
    s.p = Eigen::MatrixXf(s.M + 1, s.N + 1);
    s.cellType = Eigen::MatrixXf(s.M, s.N);
    for (int i = 0; i <= s.M; ++i)
        for (int j = 0; j < s.N; ++j) s.u(i, j) = float(100 * i + j);
    const std::vector<double> ru = hipRowMajor(s.u);
    bool order = ru.size() == size_t(s.M + 1) * s.N;
    for (int i = 0; i <= s.M && order; ++i)
        for (int j = 0; j < s.N; ++j) order = order && ru[size_t(i) * s.N + j] == 100.0 * i + j;
    expect(order, "row-major export of a column-major field");
    Eigen::MatrixXf back(s.M + 1, s.N);
    hipFromRowMajor(ru, back);
    bool same = true;
    for (int i = 0; i <= s.M; ++i)
        for (int j = 0; j < s.N; ++j) same = same && back(i, j) == s.u(i, j);
    expect(same, "row-major import");

    for (int i = 0; i < s.M; ++i)
        for (int j = 0; j < s.N; ++j) s.cellType(i, j) = (i == 0 || i == s.M - 1) ? 1.0f : 0.0f;
    for (int i = 0; i <= s.M; ++i)
        for (int j = 0; j < s.N; ++j) s.u(i, j) = 0.0f;
    s.hipPressureAlgorithm = 1;
    s.momentumSolverType = 1;
    try {
        s.hipBackendInit();
        expect(expectDevice, "hipBackendInit opened a device session");
        int itP = -1;
        float err = 0.0f;
        for (int k = 0; k < 3; ++k) err = s.calculateStepHip(itP);
        expect(std::isfinite(err) && err > 0.0f && itP >= 0, "three outer iterations on the device");
        const PlaneMetrics m = samplePlaneAtXHip(s, 0.0f);
        expect(m.valid && m.flowArea > 0.0f, "inlet plane on the device");
        s.hipBackendSync();
        expect(s.u(2, 0) == 0.1f, "inlet velocity synced back");
        s.hipBackendShutdown();
    } catch (const std::runtime_error& e) {
        std::printf("error: %s\n", e.what());
        const std::string msg = e.what();
        expect(!expectDevice && (msg.find("no HIP device") != std::string::npos ||
                                 msg.find("gfx950") != std::string::npos),
               "without a gfx950 device hipBackendInit reports CP_E_NODEV's message");
    }
    return fails ? 1 : 0;
}
