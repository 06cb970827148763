// hip_backend.cpp — binding of the MI355X SIMPLE backend into the
// reference's class SIMPLE.
//
// A maintainer adds this file to the reference as
// Utilities/solvers/gpu/hip_backend.cpp, adds the members listed in
// hip_backend.h to class SIMPLE (SIMPLE.h), and builds with
//     -DUSE_COLDPLATE_HIP -I<repo>/include -I<repo>/coldplate-solver_amd/binding
//     -L<repo>/coldplate-solver_amd/lib -lcoldplate_hip
// The backend is selected by pressureSolverType == 5 (SIMPLE.h:145):
// calculateStep (iterations.cpp:82) then delegates the whole outer
// iteration to cp_step, and the fields stay on the device between
// iterations. momentumSolverType keeps its meaning (SIMPLE.h:172: 0 =
// explicit, 1 = implicit) and picks the device's momentum algorithm;
// hipPressureAlgorithm picks the device's pressure algorithm in the
// reference's numbering (0 = SOR, 1 = PCG).
#include "SIMPLE.h"

#include <stdexcept>
#include <string>
#include <vector>

#include "coldplate_defaults.h"
#include "coldplate_hip.h"
#include "hip_backend.h"

namespace {

cp_ctx* ctxOf(const SIMPLE& s) { return static_cast<cp_ctx*>(s.hipCtx); }

void check(int rc, const void* handle) {
    if (rc != CP_OK) throw std::runtime_error(std::string("coldplate HIP backend: ") + cp_last_error(handle));
}

}  // namespace

// Eigen matrices are column-major; the C ABI takes row-major (i, j) arrays.
template <class Mat>
std::vector<double> hipRowMajor(const Mat& a) {
    std::vector<double> out(size_t(a.rows()) * size_t(a.cols()));
    for (int i = 0; i < a.rows(); ++i)
        for (int j = 0; j < a.cols(); ++j) out[size_t(i) * a.cols() + j] = a(i, j);
    return out;
}

template <class Mat>
void hipFromRowMajor(const std::vector<double>& in, Mat& a) {
    for (int i = 0; i < a.rows(); ++i)
        for (int j = 0; j < a.cols(); ++j) a(i, j) = float(in[size_t(i) * a.cols() + j]);
}

// The solve's parameters from the SIMPLE members of the same meaning
// (SIMPLE.h:68-358). rho and eta are the object's own values.
cp_params hipParamsOf(const SIMPLE& s) {
    cp_params p;
    cp_fill_reference_defaults(&p);
    p.M = s.M;
    p.N = s.N;
    p.hx = s.hx;
    p.hy = s.hy;
    p.rho = s.rho;
    p.mu = s.eta;
    p.U_in = s.targetVel;
    p.enableInletRamp = s.enableInletRamp;
    p.rampSteps = s.rampSteps;
    p.uvAlpha = s.uvAlpha;
    p.pAlpha = s.pAlpha;
    p.convectionScheme = s.convectionScheme;
    p.enableTwoPointFiveD = s.enableTwoPointFiveD;
    p.enableConvectionScaling = s.enableConvectionScaling;
    p.Ht = s.Ht_channel;
    p.sinkMultiplier = s.twoPointFiveDSinkMultiplier;
    p.alphaMax = s.brinkmanAlphaMax;  // computed at start-up as eta / K_min (SIMPLE.cpp:65-66)
    p.KMin = s.brinkmanKMin;
    p.brinkmanQ = s.brinkmanQ;
    p.enablePseudoTime = s.enablePseudoTimeStepping;
    p.useLocalPseudoTime = s.useLocalPseudoTime;
    p.timeStepMultiplier = s.timeStepMultiplier;
    p.minPseudoSpeed = s.minPseudoSpeed;
    p.enableCflRamp = s.enableCflRamp;
    p.pseudoCFLInitial = s.pseudoCFLInitial;
    p.pseudoCFLMax = s.pseudoCFLMax;
    p.cflRampStartRes = s.cflRampStartRes;
    p.cflRampExponent = s.cflRampExponent;
    p.cflRampSmooth = s.cflRampSmooth;
    p.enableSER = s.enableSER;
    p.serMinIter = s.serMinIter;
    p.serUseMaxResid = s.serUseMaxResid;
    p.serCFLMin = s.serCFLMin;
    p.serCFLMax = s.serCFLMax;
    p.serSmooth = s.serSmooth;
    p.serMaxIncrease = s.serMaxIncrease;
    p.serMaxDecrease = s.serMaxDecrease;
    p.enableLineSearch = s.enableLineSearch;
    p.lsMaxTries = s.lsMaxTries;
    p.lsAlphaMin = s.lsAlphaMin;
    p.lsAlphaReduce = s.lsAlphaReduce;
    p.lsResidIncreaseTol = s.lsResidIncreaseTol;
    p.momentumTol = s.momentumTol;
    p.maxMomentumIter = s.maxMomentumIter;
    p.pTol = s.pTol;
    p.maxPressureIter = s.maxPressureIter;
    p.epsilon = s.epsilon;
    p.maxIterations = s.maxIterations;
    p.usePressureDropConvergence = s.usePressureDropConvergence;
    p.dpConvergenceWindow = s.dpConvergenceWindow;
    p.dpConvergencePercent = s.dpConvergencePercent;
    p.usePressureDropSlopeGate = s.usePressureDropSlopeGate;
    p.dpSlopeWindowIters = s.dpSlopeWindowIters;
    p.dpSlopeMaxDelta = s.dpSlopeMaxDelta;
    p.xPlaneCoreInlet = s.xPlaneCoreInlet;
    p.xPlaneCoreOutlet = s.xPlaneCoreOutlet;
    p.pressureSolverType = s.hipPressureAlgorithm;
    p.momentumSolverType = s.momentumSolverType;
    p.sorOmega = s.sorOmega;
    p.precision = s.hipPrecision;
    return p;
}

// After buildAlphaFromDensity() (SIMPLE.cpp:25): parameters, geometry and the
// current (initial or restart) fields go to the device once.
void SIMPLE::hipBackendInit() {
    if (hipCtx) hipBackendShutdown();
    const cp_params prm = hipParamsOf(*this);
    const std::vector<double> ct = hipRowMajor(cellType);
    cp_ctx* ctx = nullptr;
    check(cp_create(&prm, ct.data(), &ctx), nullptr);
    hipCtx = ctx;
    const std::vector<double> u0 = hipRowMajor(u), v0 = hipRowMajor(v), p0 = hipRowMajor(p);
    check(cp_set_fields(ctx, u0.data(), v0.data(), p0.data()), ctx);
}

// The body of calculateStep (iterations.cpp:82-526) when pressureSolverType
// == 5: one outer iteration on the device, including the inlet ramp and the
// SER / line-search / CFL-ramp updates of SIMPLE.cpp:321-327 (the caller
// skips its own). The residuals and the CFL come back; the fields stay.
float SIMPLE::calculateStepHip(int& pressureIterations) {
    cp_step_stats st;
    check(cp_step(ctxOf(*this), &st), hipCtx);
    if (st.health != 0) throw std::runtime_error("coldplate HIP backend: non-finite field (CP_E_DIVERGED)");
    pressureIterations = st.itP;
    residU = float(st.residU);
    residV = float(st.residV);
    residMass = float(st.residMass);
    pseudoCFL = float(st.pseudoCFL);
    return float(st.error);
}

// samplePlaneAtX (postprocessing.cpp:97-231) on the device fields, for the
// monitoring of SIMPLE.cpp:338-355 and 449-457 without a field copy.
PlaneMetrics samplePlaneAtXHip(const SIMPLE& solver, float xPhysical) {
    cp_plane_metrics m;
    check(cp_plane(ctxOf(solver), xPhysical, 1.0 - 1e-6, &m), solver.hipCtx);
    PlaneMetrics out;
    out.flowArea = float(m.flowArea);
    out.massFlux = float(m.massFlux);
    out.avgStatic = float(m.avgStatic);
    out.avgDynamic = float(m.avgDynamic);
    out.avgTotal = float(m.avgTotal);
    out.valid = m.valid != 0;
    return out;
}

// Before saveAll() (SIMPLE.cpp:433, 443): the device fields into u, v, p.
void SIMPLE::hipBackendSync() {
    std::vector<double> hu(size_t(M + 1) * N), hv(size_t(M) * (N + 1)), hp(size_t(M + 1) * (N + 1));
    check(cp_get_fields(ctxOf(*this), hu.data(), hv.data(), hp.data()), hipCtx);
    hipFromRowMajor(hu, u);
    hipFromRowMajor(hv, v);
    hipFromRowMajor(hp, p);
}

// In ~SIMPLE.
void SIMPLE::hipBackendShutdown() {
    cp_destroy(ctxOf(*this));
    hipCtx = nullptr;
}
