// Minimal stand-in for the reference's SIMPLE.h, for compiling the binding
// (coldplate-solver_amd/binding/hip_backend.cpp) without Eigen: it declares
// only the members the binding touches, with the names and types of
// SIMPLE.h (checked against the real header when it is present), plus the
// members hip_backend.h asks a maintainer to add. Eigen::MatrixXf is
// replaced by a column-major float matrix with the same element access.
#pragma once

#include <cstddef>
#include <vector>

namespace Eigen {
class MatrixXf {
public:
    MatrixXf() = default;
    MatrixXf(int r, int c) : r_(r), c_(c), d_(size_t(r) * size_t(c), 0.0f) {}
    int rows() const { return r_; }
    int cols() const { return c_; }
    float& operator()(int i, int j) { return d_[size_t(j) * r_ + i]; }
    float operator()(int i, int j) const { return d_[size_t(j) * r_ + i]; }

private:
    int r_ = 0, c_ = 0;
    std::vector<float> d_;
};
}  // namespace Eigen

struct PlaneMetrics {  // SIMPLE.h:679-686
    float flowArea = 0.0f;
    float massFlux = 0.0f;
    float avgStatic = 0.0f;
    float avgDynamic = 0.0f;
    float avgTotal = 0.0f;
    bool valid = false;
};

class SIMPLE {
public:
    // SIMPLE.h:68-358, the reference defaults
    float epsilon = 1e-5f;
    int maxIterations = 10000;
    float xPlaneCoreInlet = 0.01f;
    float xPlaneCoreOutlet = 0.050f;
    bool usePressureDropConvergence = true;
    int dpConvergenceWindow = 500;
    float dpConvergencePercent = 1.0f;
    bool usePressureDropSlopeGate = false;
    int dpSlopeWindowIters = 1000;
    float dpSlopeMaxDelta = 10.0f;
    int pressureSolverType = 4;
    int maxPressureIter = 20;
    float pTol = 1e-4f;
    float sorOmega = 0.0f;
    int momentumSolverType = 1;
    float momentumTol = 1e-5f;
    int maxMomentumIter = 200;
    int convectionScheme = 0;
    bool enableTwoPointFiveD = false;
    bool enableConvectionScaling = false;
    float Ht_channel = 0.0f;
    float twoPointFiveDSinkMultiplier = 4.8f;
    float brinkmanKMin = 1e-14f;
    float brinkmanAlphaMax = 0.0f;
    float brinkmanQ = 1.0f;
    float timeStepMultiplier = 0.01f;
    bool enablePseudoTimeStepping = false;
    float pseudoCFL = 0.1f;
    float minPseudoSpeed = 0.05f;
    bool useLocalPseudoTime = true;
    bool enableSER = true;
    float serCFLMin = 0.1f;
    float serCFLMax = 5.0f;
    float serSmooth = 0.5f;
    int serMinIter = 10;
    float serMaxIncrease = 1.1f;
    float serMaxDecrease = 0.7f;
    bool serUseMaxResid = false;
    bool enableLineSearch = true;
    float lsAlphaMin = 0.1f;
    float lsAlphaReduce = 0.3f;
    int lsMaxTries = 4;
    float lsResidIncreaseTol = 1.01f;
    bool enableCflRamp = true;
    float pseudoCFLInitial = 0.1f;
    float pseudoCFLMax = 30.0f;
    float cflRampStartRes = 1e-4f;
    float cflRampExponent = 0.8f;
    float cflRampSmooth = 0.1f;
    int M = 0;
    int N = 0;
    float hy = 0.0f;
    float hx = 0.0f;
    float rho = 997.0f;
    float eta = 0.00089f;
    float targetVel = 0.0f;
    bool enableInletRamp = false;
    int rampSteps = 1000;
    float uvAlpha = 0.7f;
    float pAlpha = 0.2f;
    float residU = 1.0f, residV = 1.0f, residMass = 1.0f;

    // SIMPLE.h:399-440
    Eigen::MatrixXf u, v, p;
    Eigen::MatrixXf cellType;

    // what hip_backend.h asks a maintainer to add
    void* hipCtx = nullptr;
    int hipPressureAlgorithm = 1;
    int hipPrecision = 0;
    void hipBackendInit();
    float calculateStepHip(int& pressureIterations);
    void hipBackendSync();
    void hipBackendShutdown();
};
