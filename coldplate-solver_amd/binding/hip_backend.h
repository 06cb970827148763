// hip_backend.h — what the reference's SIMPLE.h gains for the MI355X
// backend (hip_backend.cpp).
//
// Members a maintainer adds to class SIMPLE (SIMPLE.h), next to
// pressureSolverType (SIMPLE.h:145):
//
//     void* hipCtx = nullptr;        // the device session (cp_ctx*) of this solve
//     int   hipPressureAlgorithm = 1; // device pressure algorithm: 0 = SOR, 1 = PCG
//     int   hipPrecision = 0;         // CP_PREC_FP64 (the drop-in) or CP_PREC_MIXED
//     void  hipBackendInit();
//     float calculateStepHip(int& pressureIterations);
//     void  hipBackendSync();
//     void  hipBackendShutdown();
//
// and, after struct PlaneMetrics (SIMPLE.h:679-686), the declarations below.
#pragma once

#include "coldplate_hip.h"

class SIMPLE;
struct PlaneMetrics;

// samplePlaneAtX (postprocessing.cpp:97-231) evaluated on the device fields.
PlaneMetrics samplePlaneAtXHip(const SIMPLE& solver, float xPhysical);

// The cp_params of a solve from the SIMPLE members of the same meaning.
cp_params hipParamsOf(const SIMPLE& s);
