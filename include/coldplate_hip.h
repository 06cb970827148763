/*
 * coldplate_hip.h — C ABI of the MI355X (gfx950) SIMPLE backend.
 *
 * The library keeps every field of a solve (staggered u/v/p, u*, v*, p',
 * dE/dN, the Brinkman alpha faces and the gamma density field) resident in
 * HBM for the whole run. The host only hands over geometry and parameters,
 * and reads fields back on request (checkpoint, result).
 *
 * Shapes follow the reference's staggered MAC layout (SIMPLE.h:393-440,
 * SIMPLE.cpp:184-201), stored row-major here:
 *     cellType / gamma : M x N
 *     u, u*, dE        : (M+1) x N
 *     v, v*, dN        : M x (N+1)
 *     p, p'            : (M+1) x (N+1)
 *
 * Error convention: every entry point returns CP_OK (0) or a negative
 * CP_E_* code; cp_last_error() gives the message. Nothing throws across
 * the boundary.
 *
 * The reference's device-library seam being replaced is
 * namespace amgcl_cuda (Utilities/solvers/gpu/cuda_solver.hpp:15-59), which
 * re-uploads a host CSR system on every call. This ABI is session-scoped
 * instead: one context owns a geometry's device state, and the
 * SIMPLE sub-steps run on it without host round trips.
 */
#ifndef COLDPLATE_HIP_H
#define COLDPLATE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CP_OK          0
#define CP_E_ARG      -1   /* invalid argument / shape */
#define CP_E_HIP      -2   /* HIP runtime error */
#define CP_E_NOMEM    -3   /* device arena exhausted */
#define CP_E_STATE    -4   /* call not valid in the current state */
#define CP_E_CAPACITY -5   /* batch has no free slot or HBM budget */
#define CP_E_NODEV    -6   /* no usable gfx950 device */
#define CP_E_DIVERGED -7   /* non-finite field detected (health word) */

/* Arithmetic modes. */
#define CP_PREC_FP64   0   /* plain FP64: the drop-in single-geometry solver */
#define CP_PREC_MIXED  1   /* FP32 stencil/Jacobi/PCG + FP64 refinement,
                              FP64 reductions and convergence checks */

/* Pressure-correction preconditioners (cp_params.pressurePrecond). */
#define CP_PRECOND_JACOBI 0  /* point Jacobi, the reference */
#define CP_PRECOND_AMG    1  /* aggregation-multigrid V-cycle */

/* Fields for cp_get_field / cp_set_field. */
#define CP_F_U      0
#define CP_F_V      1
#define CP_F_P      2
#define CP_F_USTAR  3
#define CP_F_VSTAR  4
#define CP_F_PSTAR  5   /* pressure correction p' */
#define CP_F_DE     6
#define CP_F_DN     7
#define CP_F_GAMMA  8   /* M x N */
#define CP_F_ALPHA  9   /* M x N */

/*
 * Solver parameters. Field names and meaning follow the public members of
 * class SIMPLE (SIMPLE.h:68-358); cp_params_default() gives the reference
 * defaults. rho and mu are parameters here (the reference hard-codes
 * 997 / 8.9e-4 at SIMPLE.h:334-335 and ignores the file values,
 * SIMPLE.cpp:48-54).
 */
typedef struct cp_params {
    /* grid (SIMPLE.h:321-326) */
    int32_t M, N;              /* cell rows (y), cell columns (x) */
    double  hx, hy;            /* cell width, height [m] */
    /* fluid (SIMPLE.h:334-335) */
    double  rho, mu;
    /* inlet (SIMPLE.h:343-346) */
    double  U_in;              /* targetVel */
    int32_t enableInletRamp, rampSteps;
    /* relaxation (SIMPLE.h:357-358) */
    double  uvAlpha, pAlpha;
    /* convection scheme (SIMPLE.h:193): 0 FOU, 1 SOU deferred correction */
    int32_t convectionScheme;
    /* 2.5D model (SIMPLE.h:205-208) */
    int32_t enableTwoPointFiveD, enableConvectionScaling;
    double  Ht, sinkMultiplier;
    /* Brinkman (SIMPLE.h:242-244). alphaMax <= 0 means mu / max(KMin,1e-20)
       (SIMPLE.cpp:65-66). */
    double  alphaMax, KMin, brinkmanQ;
    /* pseudo-transient continuation (SIMPLE.h:260-308) */
    int32_t enablePseudoTime, useLocalPseudoTime;
    double  timeStepMultiplier, minPseudoSpeed;
    int32_t enableCflRamp;
    double  pseudoCFLInitial, pseudoCFLMax, cflRampStartRes, cflRampExponent, cflRampSmooth;
    int32_t enableSER, serMinIter, serUseMaxResid;
    double  serCFLMin, serCFLMax, serSmooth, serMaxIncrease, serMaxDecrease;
    int32_t enableLineSearch, lsMaxTries;
    double  lsAlphaMin, lsAlphaReduce, lsResidIncreaseTol;
    /* inner solvers (SIMPLE.h:152-153, 176-177) */
    double  momentumTol;  int32_t maxMomentumIter;
    double  pTol;         int32_t maxPressureIter;
    /* outer loop (SIMPLE.h:68-69) */
    double  epsilon;      int32_t maxIterations;
    /* arithmetic mode, CP_PREC_* */
    int32_t precision;
    /* CP_PREC_MIXED only: FP64 refinement cycles of the pressure
       correction (each one FP32 PCG of at most maxPressureIter iterations),
       and fixed Jacobi sweeps per momentum solve (0 = reference stop test). */
    int32_t refineCycles;
    int32_t fixedMomentumSweeps;
    /* Columns of the mass residual residMass: 0 = j in [1, N-1] as the
       reference (pressure_solver.cpp:471-483); 1 = j in [1, N-2], the p'
       unknowns only. Column N-1 carries no p' unknown, so its imbalance is
       never corrected and floors residMass (1.655e-6 kg/s/m on the 200x50
       straight channel); scope 1 lets the stop test see convergence. */
    int32_t massResidualScope;
    /* Preconditioner of the pressure-correction CG: 0 = point Jacobi, the
       reference (pressure_solver.cpp:121-184); 1 = one V-cycle of
       aggregation multigrid: 2x2 piecewise-constant aggregates with Galerkin
       coarse operators (again 5-point), two damped-Jacobi sweeps (weight
       mgOmega) before and after the coarse correction, the correction
       scaled by mgScale (unsmoothed aggregation under-corrects), and a dense
       solve on the coarsest level (<= 64 unknowns). The V-cycle is a
       symmetric positive definite operator, so CG and its stop tests are
       unchanged. */
    int32_t pressurePrecond;
    double  mgOmega, mgScale;
    /* Pressure-drop convergence (SIMPLE.h:88-102, SIMPLE.cpp:356-409): after
       iteration rampSteps, the core-plane total-pressure drop between x =
       xPlaneCoreInlet and x = xPlaneCoreOutlet enters a moving window of
       dpConvergenceWindow iterations; the run stops when the window's
       standard deviation and its latest value's distance from the mean are
       both below dpConvergencePercent % of |mean| (at least 1), and, with
       the slope gate, the least-squares slope of the last dpSlopeWindowIters
       drops is at most dpSlopeMaxDelta / dpSlopeWindowIters. Windows hold
       at most CP_DP_WINDOW_MAX iterations. */
    int32_t usePressureDropConvergence, dpConvergenceWindow;
    double  dpConvergencePercent;
    int32_t usePressureDropSlopeGate, dpSlopeWindowIters;
    double  dpSlopeMaxDelta;
    double  xPlaneCoreInlet, xPlaneCoreOutlet;
    /* The reference's backend seams (SIMPLE.h:145, 156, 172).
       pressureSolverType 0 = red-black SOR on p' (pressure_solver.cpp:
       809-909): at most maxPressureIter red+black sweeps, stopping once a
       sweep pair changes no p' by pTol or more; sorOmega outside (0, 2)
       selects 2 / (1 + sqrt(1 - cos^2(pi / max(M, N)))) clamped to
       [1, 1.95] (pressure_solver.cpp:811-815). pressureSolverType 1-4 =
       the CG backend (the reference's 1 = Jacobi-preconditioned CG, 2-4 =
       its other solvers of the same system; pressurePrecond picks the
       preconditioner here).
       momentumSolverType 0 = one explicit point update of u* and v* from
       the previous iterate (iterations.cpp:176-441); 1 = implicit assembly
       and Jacobi (momentum_solver.cpp:164-339). */
    int32_t pressureSolverType, momentumSolverType;
    double  sorOmega;
} cp_params;

#define CP_DP_WINDOW_MAX 4096

/* Statistics of one outer iteration (one pass of SIMPLE.cpp:299-439). */
typedef struct cp_step_stats {
    double  error;       /* max(residU, residV, residMass), iterations.cpp:525 */
    double  residU, residV, residMass;
    double  pseudoCFL;   /* after SER / line search / ramp */
    int32_t itU, itV;    /* Jacobi sweeps of the U / V momentum solves */
    int32_t itP;         /* PCG iterations (return value of solvePCG_CSR) */
    int32_t iter;        /* outer iterations completed */
    int32_t converged;   /* error <= epsilon */
    int32_t health;      /* 0 ok, else CP_E_DIVERGED */
} cp_step_stats;

/* PlaneMetrics (SIMPLE.h:679-686). */
typedef struct cp_plane_metrics {
    double  flowArea, massFlux, avgStatic, avgDynamic, avgTotal;
    int32_t valid;
} cp_plane_metrics;

/* ------------------------------------------------------------------ */
/* parameters and diagnostics                                          */
/* ------------------------------------------------------------------ */

/* Fill p with the reference defaults of class SIMPLE (SIMPLE.h:68-358). */
int cp_params_default(cp_params* p);

/* Message of the last error of ctx (NULL: the library-wide last error). */
const char* cp_last_error(const void* ctx);

/* Number of usable gfx950 devices (0 without a GPU); devices of other
   architectures are not counted. cp_create / cp_batch_create return
   CP_E_NODEV when the current device is not a gfx950. */
int cp_device_count(void);

/* ------------------------------------------------------------------ */
/* single-geometry context: the drop-in backend                        */
/* ------------------------------------------------------------------ */
typedef struct cp_ctx cp_ctx;

/*
 * Create a solve on the current device from parameters and the cellType
 * field (M x N, 0 = fluid, 1 = solid, intermediate = buffer).
 * Replaces SIMPLE::SIMPLE's loadParameters / initializeMemory /
 * loadTopology / loadDensityField / buildAlphaFromDensity
 * (SIMPLE.cpp:18-29, 31-233, 517-630) for the device-resident state.
 */
int cp_create(const cp_params* prm, const double* cellType, cp_ctx** out);
int cp_destroy(cp_ctx* ctx);

/* Restart / checkpoint (SIMPLE.cpp:203-232, output.cpp:73-82). Any pointer
   may be NULL to skip that field. cp_set_fields also sets u* = uOld = u and
   v* = vOld = v as the reference restart does. */
int cp_set_fields(cp_ctx* ctx, const double* u, const double* v, const double* p);
int cp_get_fields(cp_ctx* ctx, double* u, double* v, double* p);
/* Read any field (CP_F_*) in its reference shape. */
int cp_get_field(cp_ctx* ctx, int which, double* out);

/* U and V implicit momentum solves followed by the velocity boundary
   conditions on (u*, v*): replaces solveUMomentumImplicit /
   solveVMomentumImplicit (momentum_solver.cpp:346-352, SIMPLE.h:705-709)
   and iterations.cpp:118-135. Fills itU/itV/residU/residV of st. */
int cp_momentum(cp_ctx* ctx, cp_step_stats* st);

/* Pressure-correction solve: replaces SIMPLE::solvePressureSystem
   (pressure_solver.cpp:462-806, pressureSolverType 1). Fills itP and
   residMass of st. */
int cp_pressure(cp_ctx* ctx, cp_step_stats* st);

/* p += pAlpha p', pressure BCs, velocity correction and velocity BCs
   (iterations.cpp:456-500, boundaries.cpp:48-134). */
int cp_correct(cp_ctx* ctx);

/* One outer iteration: inlet ramp, calculateStep (iterations.cpp:82-526),
   then the SER / line-search / CFL-ramp updates (SIMPLE.cpp:303-327). */
int cp_step(cp_ctx* ctx, cp_step_stats* st);

/* runIterations without I/O (SIMPLE.cpp:238-439): steps until
   error <= epsilon or maxIterations. The pressure-drop-window exit is off.
   check_every: outer iterations between host convergence checks
   (<= 0: choose automatically). */
int cp_run(cp_ctx* ctx, int check_every, cp_step_stats* last);

/* samplePlaneAtX (postprocessing.cpp:97-231) at physical x. A row is kept
   when both bracketing cells have gamma >= gamma_min; the reference uses
   1 - 1e-6 (postprocessing.cpp:145-149). */
int cp_plane(cp_ctx* ctx, double x, double gamma_min, cp_plane_metrics* out);

/* Wait for all device work of ctx. */
int cp_sync(cp_ctx* ctx);

/* ------------------------------------------------------------------ */
/* batched / served form: many geometries packed in one launch         */
/* ------------------------------------------------------------------ */
typedef struct cp_batch cp_batch;

/* Slot status returned by cp_batch_poll. */
#define CP_SLOT_FREE      0
#define CP_SLOT_RUNNING   1
#define CP_SLOT_CONVERGED 2
#define CP_SLOT_MAXITER   3
#define CP_SLOT_FAILED    4

typedef struct cp_slot_info {
    int32_t status;      /* CP_SLOT_* */
    int32_t iter;        /* outer iterations done */
    double  error, residU, residV, residMass;
    int32_t health;
    int32_t M, N;
    int32_t itU, itV, itP;  /* inner iterations of the last outer iteration */
    double  pseudoCFL;
} cp_slot_info;

/*
 * A batch owns one HBM arena of hbm_budget bytes and at most max_slots
 * concurrent geometries, all in one precision mode. Each cp_batch_run
 * step is one persistent launch per SIMPLE sub-step over every running
 * slot; converged slots are skipped on the device.
 */
int cp_batch_create(int precision, size_t hbm_budget, int max_slots, cp_batch** out);
int cp_batch_destroy(cp_batch* b);
/* Bytes of arena a geometry of M x N cells needs in this batch. */
size_t cp_batch_footprint(const cp_batch* b, int M, int N);
/* Admit a geometry; CP_E_CAPACITY when no slot or arena space is free. */
int cp_batch_admit(cp_batch* b, const cp_params* prm, const double* cellType, int* slot);
/* Run `steps` outer iterations of every running slot. */
int cp_batch_run(cp_batch* b, int steps);
/* Per-slot status; info has max_slots entries. */
int cp_batch_poll(cp_batch* b, cp_slot_info* info);
/* Read fields / plane metrics of a slot (same shapes as cp_get_fields). */
int cp_batch_get_fields(cp_batch* b, int slot, double* u, double* v, double* p);
/* Restart a slot from given fields, as cp_set_fields does for a context
   (SIMPLE.cpp:203-232): any pointer may be NULL; u* = u and v* = v. Only
   that slot's arrays change. Between cp_batch_run calls only. */
int cp_batch_set_fields(cp_batch* b, int slot, const double* u, const double* v, const double* p);
int cp_batch_plane(cp_batch* b, int slot, double x, double gamma_min, cp_plane_metrics* out);
/* Release a slot's arena space. */
int cp_batch_evict(cp_batch* b, int slot);

/* Launch shape of the batch (no counterpart in the reference): tile_rows R
   in {1, 2, 4, 8} gives 64 x 4R-node tiles for every slot (0 = keep; the
   default is 4), wg_per_cu the workgroups per CU of the replayed launches
   (0 = keep; default 16). Changing R retiles the admitted slots; results of
   a slot depend on R only through the order of its partial sums.
   cp_batch_autotune chooses both by timing. */
int cp_batch_set_tiling(cp_batch* b, int tile_rows, int wg_per_cu);

/* 1 (default): cp_batch_run replays the outer iteration as a captured
   hipGraph; 0: it launches the same kernel sequence directly (no
   counterpart in the reference; used for tracing tools that do not follow
   graph replays). */
int cp_batch_set_graph(cp_batch* b, int use_graph);

/* Rows per wave R of the finest multigrid level's 60 x 4R tiles (no
   counterpart in the reference; the coarser levels use 8 in mixed precision
   and 2 in FP64): 2, 4 or 8 (4 and 8 in mixed precision only); 0 derives it
   from the batch tile height. Every slot is retiled. */
int cp_batch_set_mg_tiling(cp_batch* b, int rows);

/* Tile order of the batch's launches (no counterpart in the reference), a
   mask: bit 0 for the node-grid kernels (momentum, pressure assembly,
   correction), bit 1 for the p' kernels (CG, refinement), bit 2 for the
   multigrid levels. A clear bit walks the tile list in one contiguous range per XCD
   (halo lines shared in that XCD's L2); a set bit walks it in launch order
   (the tiles in flight cover one stretch of memory, faster for DRAM once
   the working set exceeds the Infinity Cache). Results do not depend on the
   order. Default 0. */
int cp_batch_set_tile_order(cp_batch* b, int order);

/* Autotuner (no counterpart in the reference): for every candidate tiling
   (rows per wave R in {1, 2, 4, 8}, finest-multigrid-level rows in
   {2, 4, 8} (2 in FP64), workgroups per CU in {4, 8, 16}) replays `steps`
   outer iterations of the admitted slots, timed with HIP events on the batch
   stream, and keeps the fastest tiling for the batch; then it times the eight
   tile orders (cp_batch_set_tile_order) with that tiling and keeps the
   fastest. The replays are real work: every candidate advances the running
   slots by steps + 1 outer iterations. Tilings differ only in the summation
   order of reductions; tile orders not at all. Outputs (any may be NULL):
   the chosen R, multigrid rows, workgroups per CU, tile order and ms per
   outer iteration. */
int cp_batch_autotune(cp_batch* b, int steps, int* rows, int* mg_rows, int* wg_per_cu, int* tile_order,
                      double* ms_per_step);

/* Kernels the measurement hooks below can time (cp_batch_probe takes a mask
   of 1 << CP_K_*). */
#define CP_K_MOM_ASM      0   /* momentum assembly with the first Jacobi sweep */
#define CP_K_PCG_Q        4   /* PCG: search direction and q = A p */
#define CP_K_MG_UP       15   /* multigrid, finest level: prolongation, post-smoothing, z */

/* Measurement hooks (no counterpart in the reference, which has no device
   timing). cp_batch_probe(b, mask) makes cp_serve / cp_serve_pool (with b
   as the first batch) time each kernel of the mask every 10 polls: a burst
   of 16 rounds of back-to-back launches between two HIP events on b's
   stream, with the GPU to itself; each round launches the kernel once per
   batch serving the queue, over that batch's running slots, so with several
   batches the rounds rotate over all of their arrays. Each kernel
   recomputes its outputs from inputs no launch of it changes, between two
   outer iterations, and the next outer iteration recomputes them before
   any reader, so a burst leaves every solve as it was. mask 0 disables.
   cp_batch_probe_read returns the bursts of the last serve, oldest first:
   kernel, launches, elapsed ms, and work units over all its launches (p'
   unknowns for CP_K_PCG_Q, and for CP_K_MG_UP those of the multigrid slots;
   (M+1) x (N+1) grid nodes for CP_K_MOM_ASM); any output but ms may be
   NULL; the return value is their count. cp_batch_probe_burst runs one burst of `launches` launches on b
   alone now, synchronises, and gives the mean ms per launch. */
int cp_batch_probe(cp_batch* b, int kernel_mask);
int cp_batch_probe_read(cp_batch* b, int* kernel, int* launches, double* ms, double* units, int cap);
int cp_batch_probe_burst(cp_batch* b, int kernel, int launches, double* ms_per_launch);

/* ------------------------------------------------------------------ */
/* request loop with continuous batching                               */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference, which solves one geometry per process
   (SIMPLE.cpp:238-512). A request is one solve: parameters, geometry, its
   arrival time, and the plane row rule of its reported pressure drop. */
typedef struct cp_request {
    const cp_params* prm;     /* precision must match the batch */
    const double* cellType;   /* M x N row-major */
    double arrival;           /* seconds after the start of cp_serve */
    double gammaMin;          /* samplePlaneAtX row rule (1 - 1e-6: reference) */
} cp_request;

typedef struct cp_result {
    int32_t status;           /* CP_SLOT_CONVERGED / _MAXITER / _FAILED */
    int32_t iter;             /* outer iterations */
    int32_t slot;             /* slot it ran in (-1: never admitted) */
    int32_t rc;               /* CP_OK; the error that rejected it; or
                                 CP_E_DIVERGED for a solve that failed on a
                                 non-finite field (status CP_SLOT_FAILED) */
    double  error;            /* final max(residU, residV, residMass) */
    double  dp;               /* full-system total-pressure drop, x = 0 .. N hx */
    double  tAdmit, tDone;    /* seconds after the start of cp_serve */
} cp_result;

/* Serve n requests with continuous batching, in real time: requests are
   admitted in arrival order as soon as they have arrived and a slot and
   arena space (the HBM footprint) are free, a request that does not fit
   waits for evictions (first come, first served); every running slot
   advances by steps_per_poll outer iterations per launch of the step
   graph, and each finished solve is read out (pressure drop) and evicted
   at once, freeing its space for the next arrivals. A request that can
   never fit, or whose parameters are invalid, gets its error code in rc.
   Returns CP_OK once every request has finished. */
int cp_serve(cp_batch* b, const cp_request* reqs, int n, int steps_per_poll, cp_result* results);

/* cp_serve over several empty batches of one precision at once, each on its
   own HIP stream and served by its own host thread (batches[0] by the
   calling thread), all taking requests from one queue, first come first
   served, into whichever batch has room. No reference counterpart (the
   reference solves one geometry per process). Kernels of different streams
   run concurrently, so one batch's latency-bound launches (per-slot
   finalisations, coarse multigrid levels) overlap another's bandwidth-bound
   ones. Timing bursts (cp_batch_probe) run with no other batch's work in
   flight. Batches are meant to be alike: a request that does not fit an
   empty batch is rejected there, as cp_serve rejects it. Times in results
   are seconds after the start of the call. Returns the first error any
   batch met; the others stop at their next poll, possibly with slots still
   admitted. */
int cp_serve_pool(cp_batch* const* batches, int nbatches, const cp_request* reqs, int n, int steps_per_poll,
                  cp_result* results);

/* ------------------------------------------------------------------ */
/* downstream conjugate heat transfer (the reference's ThermalSolver)   */
/* ------------------------------------------------------------------ */
/* Parameters of the 3D conjugate-heat-transfer solve; names and defaults
   follow struct Params of ThermalSolver/thermal_solver.cpp:142-189. The
   grid is Nx x Ny cells in the plane and nz_solid base-plate layers (H_b
   thick) under nz_fluid channel layers (H_t thick). */
typedef struct cp_thermal_params {
    int32_t Nx, Ny;               /* cells in x (columns) and y (rows) */
    int32_t nz_solid, nz_fluid;   /* 2, 7 */
    double  dx, dy;               /* cell width, height [m] */
    double  H_b, H_t;             /* base plate and channel heights [m]: 5e-4, 4e-3 */
    double  k_s, k_f;             /* solid / fluid conductivity [W/m K]: 400, 0.6 */
    double  rho, Cp;              /* fluid density, heat capacity: 998, 4180 */
    double  T_inlet;              /* 30 */
    double  q0;                   /* heat flux into the base [W/m^2] (thermal_params.txt) */
    int32_t k_interp_mode;        /* 0 linear, 1 RAMP (default), 2 harmonic */
    double  qk;                   /* RAMP convexity: 1 */
    double  rtol;                 /* relative residual of the linear solve: 1e-6 */
    int32_t maxiter;              /* 1000 */
    int32_t timing;               /* nonzero: fill the timing fields of cp_thermal_stats (no reference
                                     counterpart; HIP events on the solve's stream) */
} cp_thermal_params;

/* Results of cp_thermal_solve: the solver's iterations and final relative
   residual, and the statistics the reference prints and writes
   (thermal_solver.cpp:556-604, thermal_output.cpp:135-205). */
typedef struct cp_thermal_stats {
    int32_t iters, converged;
    double  relres;
    double  T_base_min, T_base_max, T_fin_min, T_fin_max, T_fluid_min, T_fluid_max;
    double  T_avg_base, T_avg_fluid, T_avg_solid, T_avg_global;
    double  Q_in, Q_out, T_out, imbalance_pct;
    /* with timing: device time of the solve (assembly to the last
       iteration, ms) and the average duration (ms) of one finest-level
       z-line sweep from a nonzero iterate, over `sweeps` such sweeps */
    double  device_ms, sweep_ms;
    int32_t sweeps;
} cp_thermal_stats;

/* Fill p with the defaults of the reference's thermal Params. */
int cp_thermal_params_default(cp_thermal_params* p);

/* Replaces ThermalSolver3D::solve (ThermalSolver/thermal_solver.cpp:421-
   604): assembles the reference's finite-volume system on the device (one
   7-point row per cell: harmonic-mean lateral conductances, series
   conductances in z, first-order upwind advection of the depth-profiled
   velocity 4z(1-z) u in fluid cells, the inlet temperature and convective
   outlet, the base heat flux), solves it in FP64 to relative residual rtol
   with BiCGSTAB preconditioned by one aggregation-multigrid V-cycle
   (2 x 2 aggregates in the plane, every layer kept, z-line Jacobi
   smoothing), and returns T and the statistics. gamma (1 = fluid), u, v:
   Ny x Nx row-major cell values, as the reference's geometry_thermal.txt
   (inverted) and u_thermal.txt / v_thermal.txt; T: (nz_solid + nz_fluid)
   x Ny x Nx, cell (i, j, k) at k Nx Ny + j Nx + i, the reference's order.
   Allocates its device arrays per call. Returns CP_E_DIVERGED when the
   residual becomes non-finite; a solve that ends at maxiter returns CP_OK
   with converged = 0. */
int cp_thermal_solve(const cp_thermal_params* p, const double* gamma, const double* u, const double* v, double* T,
                     cp_thermal_stats* st);

/* ------------------------------------------------------------------ */
/* a solve that allocates no device memory                             */
/* ------------------------------------------------------------------ */
/* Device scratch of cp_run_noalloc: a static array of this many bytes that
   every device the library runs on holds until the process ends. */
#define CP_NOALLOC_SCRATCH_BYTES (32u << 20)

/* No counterpart in the reference. For callers that cannot allocate device
   memory (tight serving loops, harnesses that forbid hipMalloc): cp_create,
   cp_set_fields, cp_run and cp_get_fields in one call that calls no
   hipMalloc, hipHostMalloc, hipFree, stream or graph creation. The solve's
   descriptor, slot state, tile lists, reduction partials and arrays live in
   the static device scratch of the current device, with the tiling cp_create
   chooses, so the results are those of cp_run on a context, bit for bit.
   u, v, p (reference shapes, row-major, host): the initial fields when
   non-NULL (else zero), overwritten with the final ones. A grid whose solve
   does not fit CP_NOALLOC_SCRATCH_BYTES gets CP_E_CAPACITY (an FP64 solve
   needs about 270 bytes per node of the (M+1) x (N+1) grid). Runs on the
   null stream and returns when the solve has ended (error <= epsilon,
   maxIterations, or a non-finite field: health in *last). Calls that
   overlap on one device share the scratch and race; the caller serialises
   them. Each call clears what it reads before writing it. */
int cp_run_noalloc(const cp_params* prm, const double* cellType, double* u, double* v, double* p,
                   cp_step_stats* last);

/* ------------------------------------------------------------------ */
/* the reference's file formats (host only, no device needed)          */
/* ------------------------------------------------------------------ */
/* Parameter file "M N hy hx U_in N_in_buffer N_out_buffer [Ht [rho nu]]"
   (SIMPLE.cpp:31-60): sets M, N, hy, hx, U_in and, when present, Ht of p;
   rho and nu columns are ignored as the reference ignores them. */
int cp_io_read_params(const char* path, cp_params* p, int* n_in_buffer, int* n_out_buffer);
/* Geometry file: M x N whitespace-separated cellType values, row-major
   (SIMPLE.cpp:517-551). Returns 1 (all fluid) when the file is missing;
   values missing at the end stay fluid. */
int cp_io_read_geometry(const char* path, int M, int N, double* cellType);
/* Whitespace-separated rows x cols matrix, row-major (the restart files
   u.txt, v.txt, p.txt: SIMPLE.cpp:203-232); CP_E_ARG when incomplete. */
int cp_io_read_matrix(const char* path, int rows, int cols, double* a);
/* Tab-separated rows, as saveMatrix (output.cpp:47-71). precision 0 writes
   the reference's 6 significant digits; 17 round-trips doubles. */
int cp_io_write_matrix(const char* path, const double* a, int rows, int cols, int precision);
/* saveAll (output.cpp:73-266) into `dir`: u/v/p.txt (staggered),
   u_full/v_full/pressure_full.txt (cell-centred), u_thermal/v_thermal/
   pressure_thermal/p_gradient.txt (buffers cropped) and fluid_results.vtk.
   gamma and alpha are the M x N fields (cp_get_field CP_F_GAMMA / _ALPHA). */
int cp_io_save_all(const char* dir, int M, int N, double hx, double hy, int n_in_buffer, int n_out_buffer,
                   const double* u, const double* v, const double* p, const double* cellType, const double* gamma,
                   const double* alpha, int precision);

/* ------------------------------------------------------------------ */
/* seeded geometry generators (cellType, M x N row-major)              */
/* ------------------------------------------------------------------ */
/* The reference ships Python generators (GeometryGeneratorV1.py ...
   GeometryGeneratorV5.py) writing geometry_fluid.txt; these produce the
   same format for the solver's workloads. */
/* Straight channel: fluid on cell rows [row0, row1], solid elsewhere. */
int cp_geom_straight(int M, int N, int row0, int row1, double* cellType);
/* Seeded network of 6-10 Manhattan/diagonal channels 8-16 cells wide with
   split/merge junctions, carved into solid. */
int cp_geom_network(int M, int N, uint64_t seed, double* cellType);
/* Seeded Gaussian random field (correlation length corr cells), mapped to
   [0,1] and tanh-projected with (beta, eta); cellType = 1 - gamma. */
int cp_geom_gamma(int M, int N, uint64_t seed, double corr, double beta, double eta, double* cellType);

#ifdef __cplusplus
}
#endif

#endif /* COLDPLATE_HIP_H */
