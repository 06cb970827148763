/*
 * coldplate_defaults.h — the reference's parameter defaults as a table.
 *
 * Values are the member initialisers of class SIMPLE (SIMPLE.h:68-358):
 * they are data, kept identical on purpose so a caller switching backends
 * gets the same behaviour. Used by cp_params_default() and by the CPU oracle.
 */
#ifndef COLDPLATE_DEFAULTS_H
#define COLDPLATE_DEFAULTS_H

#include <string.h>
#include "coldplate_hip.h"

static inline void cp_fill_reference_defaults(cp_params* p) {
    memset(p, 0, sizeof(*p));
    p->epsilon = 1e-5;            p->maxIterations = 10000;       /* SIMPLE.h:68-69 */
    p->maxPressureIter = 20;      p->pTol = 1e-4;                 /* SIMPLE.h:152-153 */
    p->momentumTol = 1e-5;        p->maxMomentumIter = 200;       /* SIMPLE.h:176-177 */
    p->convectionScheme = 0;                                      /* SIMPLE.h:193 */
    p->enableTwoPointFiveD = 0;   p->enableConvectionScaling = 0; /* SIMPLE.h:205-206 */
    p->Ht = 0.0;                  p->sinkMultiplier = 4.8;        /* SIMPLE.h:207-208 */
    p->KMin = 1e-14;              p->alphaMax = 0.0;  p->brinkmanQ = 1.0;  /* SIMPLE.h:242-244 */
    p->timeStepMultiplier = 0.01; p->enablePseudoTime = 0;        /* SIMPLE.h:260-262 */
    p->minPseudoSpeed = 0.05;     p->useLocalPseudoTime = 1;      /* SIMPLE.h:264-265 */
    p->enableSER = 1;  p->serCFLMin = 0.1;  p->serCFLMax = 5.0;   /* SIMPLE.h:278-280 */
    p->serSmooth = 0.5; p->serMinIter = 10;                        /* SIMPLE.h:282-283 */
    p->serMaxIncrease = 1.1; p->serMaxDecrease = 0.7; p->serUseMaxResid = 0; /* SIMPLE.h:284-286 */
    p->enableLineSearch = 1; p->lsAlphaMin = 0.1; p->lsAlphaReduce = 0.3;    /* SIMPLE.h:291-293 */
    p->lsMaxTries = 4; p->lsResidIncreaseTol = 1.01;               /* SIMPLE.h:294-295 */
    p->enableCflRamp = 1; p->pseudoCFLInitial = 0.1; p->pseudoCFLMax = 30.0; /* SIMPLE.h:303-305 */
    p->cflRampStartRes = 1e-4; p->cflRampExponent = 0.8; p->cflRampSmooth = 0.1; /* SIMPLE.h:306-308 */
    p->rho = 997.0;  p->mu = 0.00089;                              /* SIMPLE.h:334-335 */
    p->U_in = 0.0;   p->enableInletRamp = 0;  p->rampSteps = 1000; /* SIMPLE.h:343-346 */
    p->uvAlpha = 0.7; p->pAlpha = 0.2;                             /* SIMPLE.h:357-358 */
    p->precision = CP_PREC_FP64;
    p->refineCycles = 2;
    p->fixedMomentumSweeps = 0;
    p->massResidualScope = 0;
    p->pressurePrecond = 0;
    p->mgOmega = 0.8;
    p->mgScale = 1.8;
    p->usePressureDropConvergence = 1; p->dpConvergenceWindow = 500;    /* SIMPLE.h:98-99 */
    p->dpConvergencePercent = 1.0;                                       /* SIMPLE.h:100 */
    p->usePressureDropSlopeGate = 0; p->dpSlopeWindowIters = 1000;      /* SIMPLE.h:101-102 */
    p->dpSlopeMaxDelta = 10.0;                                           /* SIMPLE.h:103 */
    p->xPlaneCoreInlet = 0.01; p->xPlaneCoreOutlet = 0.050;             /* SIMPLE.h:88-89 */
    p->pressureSolverType = 1;  /* the CG backend (the reference's default 4 is AMG-CG on a GPU) */
    p->momentumSolverType = 1;                                           /* SIMPLE.h:172 */
    p->sorOmega = 0.0;                                                   /* SIMPLE.h:156 */
}

/* Defaults of the thermal solve: struct Params of
   ThermalSolver/thermal_solver.cpp:142-189. Grid sizes, spacings and the
   heat flux come from the caller (the reference reads them from
   thermal_params.txt and the geometry file). */
static inline void cp_fill_thermal_defaults(cp_thermal_params* p) {
    memset(p, 0, sizeof(*p));
    p->nz_solid = 2; p->nz_fluid = 7;              /* :147-148 */
    p->H_b = 0.0005; p->H_t = 0.004;               /* :152-153 */
    p->k_s = 400.0; p->k_f = 0.6;                  /* :156-157 */
    p->rho = 998.0; p->Cp = 4180.0;                /* :158-159 */
    p->T_inlet = 30.0;                             /* :160 */
    p->rtol = 1e-6; p->maxiter = 1000;             /* :163-164 */
    p->k_interp_mode = 1; p->qk = 1.0;             /* :179-180 */
}

#endif /* COLDPLATE_DEFAULTS_H */
