// cp_device.h — device-side data layout shared by the kernels and the host
// engine of the gfx950 SIMPLE backend.
//
// Every geometry ("slot") lives in one arena region of the batch. All of its
// fields are stored on one node grid of (M+1) rows x pitch columns,
// row-major, pitch >= N+1 and a multiple of 32 elements. A node (i, j) holds
//   u(i,j)  [valid for i<=M,   j<=N-1]     v(i,j)  [i<=M-1, j<=N]
//   p(i,j), p'(i,j) [i<=M, j<=N]
//   gam(i,j) = gamma(i-1, j-1)  (cell-centred fields moved onto the p index)
// so the staggered stencils of the reference (SIMPLE.h:17-31) become fixed
// offsets on one grid. Kernels cover the grid with 64 x 4R node tiles; one
// workgroup of 256 threads (4 wave64) processes one tile per visit, each wave
// a 64-column strip of R rows, marching down the rows.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpdev {

constexpr int kTileW = 64;
constexpr int kTileH = 4;
constexpr int kBlock = kTileW * kTileH;
// Multigrid level tiles: the 64 lanes of a wave span kMgTileW columns of the
// level plus a two-column halo on each side (the two smoothing sweeps and the
// residual reach two nodes out).
constexpr int kMgTileW = kTileW - 4;
// A multigrid tile is R rows x kMgTileCols columns: the block's kTileH waves
// side by side, wave w on columns [J0 + w kMgTileW, J0 + (w+1) kMgTileW), so
// that the block reads each row as one contiguous run.
constexpr int kMgTileCols = kTileH * kMgTileW;
constexpr int kNumPartials = 5;  // per-tile reduction partials (doubles)
constexpr int kNumXCD = 8;       // accelerator complex dies of an MI355X
constexpr int kMaxLevels = 12;   // aggregation-multigrid levels (incl. the finest)
constexpr int kCoarseMax = 64;   // unknowns of the coarsest level (dense inverse)
// The coarse end of the hierarchy runs in one workgroup per slot, in LDS:
// from the first level whose remaining levels hold at most this many
// unknowns in total (per array) down to the coarsest and back.
constexpr int kTailCapF32 = 3072;
constexpr int kTailCapF64 = 1536;

// slot status
constexpr int ST_RUNNING = 1;
constexpr int ST_CONVERGED = 2;
constexpr int ST_MAXITER = 3;
constexpr int ST_FAILED = 4;

// Solver-typed arrays (T = double in FP64 mode, float in mixed mode).
// Momentum system (U or V): the relaxed diagonal aP, the east and north face
// fluxes Fe, Fn of each row (its neighbour coefficients are De + max(0, -Fe),
// De + max(0, Fe of the west row), Dn + max(0, -Fn), Dn + max(0, Fn of the
// south row): a face's flux is shared by the rows on either side of it), the
// initial residual r0 and two buffers of the Jacobi correction delta.
enum MomArr { M_AP = 0, M_FE, M_FN, M_R0, M_D0, M_D1, M_COUNT };
// P_Z: preconditioned residual z = M^-1 r (multigrid); P_MGX: the finest
// level's pre-smoothed iterate. P_X stays last: FP64 aliases it to p'.
enum PresArr { P_CE = 0, P_CN, P_AP, P_R, P_PK0, P_PK1, P_Q, P_Z, P_MGX, P_X, P_COUNT };

// One level of the aggregation-multigrid hierarchy of a slot. Unknown (I, J),
// 0 <= I < H, 0 <= J < W, is stored at row I+1, column J+1 of an array with
// `pitch` columns; the finest level is the p' grid itself (p'(i,j) at node
// (i,j), i = I+1, j = J+1, on the slot's node grid). Level l+1 aggregates
// 2x2 blocks of level l. Its tiles are 64 x 4R unknowns from (0,0).
struct MgLevel {
    int H, W, pitch;
    int tile0, ntiles, tilesX;
    void* aP;
    void* cE;
    void* cN;
    void* b;  // right-hand side (finest: the PCG residual r)
    void* x;  // iterate after pre-smoothing
    void* z;  // V-cycle result (finest: P_Z)
};

// Read-only description of one slot; uploaded when a slot is admitted.
struct alignas(64) SlotDesc {
    int M, N, pitch;
    int tile0, ntiles, tilesX;
    // the p' unknowns' own tile list (kernels that touch only p' unknowns):
    // 64 Cp x 4Rp tiles from node (1, 0), covering rows 1 .. M-1, columns
    // 0 .. N-2, with Rp = pRows(mixed, tileRows) and Cp = pCols(mixed)
    int ptile0, pntiles, ptilesX;
    int tileRows;                // rows per wave R: a tile is 64 x 4R nodes
    int nU, nV;                  // momentum DOF counts (Jacobi stop test)
    int convScheme, pseudo, localPseudo;
    int maxMomIter, maxPresIter, refineCycles, massScope;
    int maxIterations, inletRamp, rampSteps;
    int enableSER, serMinIter, serUseMax, enableLS, lsMaxTries, enableRamp;
    int mixed;
    int dE32;                    // dE, dN hold floats (mixed-precision implicit momentum, which forms
                                 // them with a single-precision reciprocal), and so do the Brinkman
                                 // faces aU, aV: their arrays' first halves
    int dpConv, dpWindow;        // pressure-drop window exit and its length
    int slopeGate, slopeWindow;  // slope gate of the window exit and its length
    double dpPercent, slopeMaxDelta, xCoreIn, xCoreOut;
    int presType, momType;       // 0: red-black SOR / explicit momentum; 1: CG / implicit + Jacobi
    double sorOmega;             // SOR over-relaxation (resolved on the host)
    int precond, nLevels;        // CP_PRECOND_*; multigrid levels (>= 2 when precond == 1)
    int mgTail;                  // first level handled by the one-workgroup tail kernel
    double mgOmega, mgScale;
    double hx, hy, rho, vol, De, Dn, convScale, sinkDiag, uvAlpha, pAlpha, hChar;
    double timeStep, minSpeed, momTol, pTol, epsilon, targetVel;
    double serCFLMin, serCFLMax, serSmooth, serMaxInc, serMaxDec;
    double lsAlphaMin, lsAlphaReduce, lsTol;
    double cflInit, cflMax, rampStart, rampExp, rampSmooth;
    // state arrays (double)
    double *u, *v, *p, *us, *vs, *pc, *dE, *dN, *aU, *aV, *gam, *bm;
    // solver arrays (T)
    void* mu[M_COUNT];
    void* mv[M_COUNT];
    void* pr[P_COUNT];
    // pressure-drop window and slope buffer (CP_DP_WINDOW_MAX each)
    double* dpHist;
    double* dpSlope;
    // multigrid hierarchy and the dense inverse of its coarsest level
    double* coarseInv;
    MgLevel lv[kMaxLevels];
};

// Mutable per-slot scalars; written only by the finalising tile of a slot
// (or by the host between launches).
struct alignas(256) SlotState {
    int status;
    int health;
    int iter;
    int jk[2], jdone[2], jiters[2];
    int pk, pdone, piters, pcycle, pitTotal;
    int itU, itV, itP;
    int stepPhase;
    double rz, alpha, beta, rr;
    double residU, residV, residMass, error;
    double inletVelocity, pseudoCFL, serResidPrev, lsAlpha;
    int lsBacktrack;
    int dpIdx, dpFilled, dpConverged, slopeHead, slopeCount;
    int sorDone;  // SOR slots: the pressure sweeps of this outer iteration have stopped
    double dpLast;  // core-plane total-pressure drop after the last iteration
    // The CG iterate of the fixed-budget mixed step with one refinement cycle
    // is formed late (kPcgQFold): the alphas of the last two updates, the
    // updates applied and those already added into x in this solve.
    double alphaPrev;
    int nUpd, nFold;
};

// Plane-sample request/result (samplePlaneAtX).
struct PlaneReq {
    int slot;
    double x;
    double gammaMin;
};
struct PlaneOut {
    double flowArea, massFlux, avgStatic, avgDynamic, avgTotal;
    int valid;
};

// Kernel arguments shared by every tile kernel.
struct KArgs {
    const SlotDesc* desc;
    SlotState* state;
    const int* tileSlot;
    const int* nTiles;
    double* partials;
    int force;  // k_pcg_q only: also run slots whose PCG has stopped (timing bursts)
    int level;  // multigrid kernels: hierarchy level of the tile list
    int color;  // SOR sweep: 0 = cells with i + j even, 1 = odd
    int order;  // tile order of the launch: 0 = one contiguous range per XCD, 1 = launch order
    int fold;   // folded steps of the fixed-budget mixed-precision step (kFold* bits)
    int pvar;   // k_pcg_q only: the slots the launch runs and their form (kPcgQ*)
};

// KArgs::pvar of a single-precision k_pcg_q launch. kPcgQAll: every slot,
// with its preconditioner's z and its first or later iteration chosen per
// slot. Otherwise one form only, which holds fewer registers (the four
// forms together take 162 VGPRs per lane at 4 x 4 values per lane, the
// multigrid form of a later iteration 122): the slots whose preconditioner
// is multigrid (kPcgQMg) or Jacobi, at the CG's first iteration
// (kPcgQFirst) or a later one. The host knows both: a batch records the
// preconditioners of its slots and launches one form per kind, and every
// running slot's CG is at the iteration the launch sequence is at.
constexpr int kPcgQAll = 0;
constexpr int kPcgQMg = 1;
constexpr int kPcgQJac = 2;
constexpr int kPcgQFirst = 4;
// With kPcgQFold (the fixed-budget step with one refinement cycle, at an even
// iteration k >= 2) the launch also adds the two updates before it to the CG
// iterate, x = (x + alpha_{k-2} p_{k-2}) + alpha_{k-1} p_{k-1}, which
// k_pcg_update leaves out in that step; k_correct adds the last one or two.
// p_{k-1} is the launch's own p_old and p_{k-2} the buffer it overwrites.
constexpr int kPcgQFold = 8;

// KArgs::fold bits, set for the fixed-budget step of a mixed-precision batch.
// kFoldMom: u* and v* of an implicit-momentum slot are not stored; the
// pressure assembly and the correction form them from u, v and the last
// Jacobi iterate's delta by the finish's rules (k_mom_finish is not
// launched), and the assembly's finalisation takes over the finish's
// residuals. kFoldPc: every slot makes one refinement cycle, so p' = 0 + x
// exactly; the correction of a CG slot reads the single-precision iterate x
// (k_refine_accum is not launched), which is formed late: k_pcg_update
// neither reads nor writes x or p_k, k_pcg_q adds two updates at a time
// (kPcgQFold) and the correction the last one or two, in the order of the
// per-iteration updates, so p' is the same bit for bit.
constexpr int kFoldMom = 1;
constexpr int kFoldPc = 2;
// Partial entries of the folded pressure assembly: the finish's max |u* - u|,
// max |v* - v| and non-finite flag follow the assembly's own two.
constexpr int kFoldFinPartial = 2;

// Finalisation phases (k_finalize), one per reducing work kernel.
enum Phase {
    PH_JAC = 0, PH_MOM_FIN, PH_PRES_ASM, PH_PCG_U, PH_REF_ACC, PH_REF_RES, PH_CORR, PH_MG_Z,
    PH_MOM_EXPL,  // explicit momentum: residuals
    PH_SOR,       // SOR: stop test after a red + black sweep pair
    PH_PCGU_MGZ,  // PH_PCG_U's stop test, then PH_MG_Z, after the V-cycle that follows the update
    PH_PCGU_REFACC,  // PH_PCG_U's stop test, then PH_REF_ACC, after the accumulation that follows the last update
};
// Partial entry of the finest up-sweep's r.z: apart from the PCG update's
// entries 0 and 1, which PH_PCGU_MGZ reads after the V-cycle.
constexpr int kMgZPartial = 2;
// Partial entry of k_pcg_q's p.q: every workgroup of the k_pcg_update that
// follows sums the slot's entries while the update's own partials (entries
// 0 and 1) are being written.
constexpr int kPcgQPartial = 3;

// Kernels of one SIMPLE step, in launch order.
enum Kern {
    K_MOM_ASM = 0,  // momentum assembly (U and V)
    K_JAC,          // one Jacobi sweep (U and V)
    K_MOM_FIN,      // u*, v* scatter + velocity BCs
    K_PRES_ASM,     // mass residual, p' coefficients, PCG start
    K_PCG_Q,        // PCG: p update + q = A p
    K_PCG_U,        // PCG: x, r update + stop test
    K_REF_ACC,      // mixed: p' += x
    K_REF_RES,      // mixed: FP64 residual of p'
    K_CORR,         // p update, velocity correction, BCs
    K_CORR_OUTER,   // K_CORR + outer-loop scalar logic
    K_MG_COARSEN,   // multigrid: Galerkin coarse operator of level a.level
    K_MG_FACTOR,    // multigrid: dense inverse of the coarsest level (one block per slot)
    K_MG_DOWN,      // multigrid: pre-smoothing + restriction at level a.level
    K_MG_TAIL,      // multigrid: levels mgTail.. down, coarsest solve, up (one block per slot)
    K_DP_MONITOR,   // pressure-drop window exit after the controller (one block per slot)
    K_MG_UP,        // multigrid: prolongation + post-smoothing at level a.level
    K_MOM_EXPL,     // explicit momentum: u*, v*, dE, dN of the unknowns
    K_MOM_EXPL_BC,  // explicit momentum: velocity BCs on (u*, v*)
    K_SOR,          // one SOR colour sweep on p' (a.color)
};

// Host-side launchers (defined next to the kernels). launchKernel enqueues
// the work kernel on `grid` workgroups and, when it reduces, the per-slot
// finalisation on `nslots` workgroups; `rows` is the batch's R (1, 2, 4 or
// 8). Non-null events are recorded right
// before and right after the work kernel (not around the finalisation).
hipError_t launchKernel(Kern k, bool f32, int rows, int grid, int nslots, hipStream_t s, const KArgs& a,
                        hipEvent_t evBefore = nullptr, hipEvent_t evAfter = nullptr);
// The work kernel alone, without its finalisation (timing bursts of an
// idempotent kernel).
hipError_t launchWorkOnly(Kern k, bool f32, int rows, int grid, hipStream_t s, const KArgs& a);
// The finalisation kernel alone (one workgroup per slot) for `phase`.
hipError_t launchFinalize(int phase, int nslots, hipStream_t s, const KArgs& a);
// Rows per wave of the finest multigrid level's tiles for a batch tile
// height. Even, so that a wave's rows hold whole 2 x 2 aggregates; 4 and 8
// only in single precision (register footprint). The two sweeps reach two
// rows out, so a wave loads R + 4 rows for its R: at 8 rows the finest
// up-sweep over 8 x 1024x512 slots takes 40.8 instead of 46.4 us per launch.
inline int mgRows(bool f32, int rows) { return (f32 && rows >= 4) ? 8 : 2; }
// Rows per wave of the coarser grid-wide levels' tiles: 8 in single
// precision (fewer, taller tiles: the 4 x 16-slot pool serves 1.8 % more
// than with 2; one batch alone is as fast), 2 in double precision.
inline int mgCoarseRows(bool f32) { return f32 ? 8 : 2; }
// Rows per wave of the p' list's tiles: 4 in single precision (measured on
// 8 x 1024x512: k_pcg_q 21.6 us with 4 against 22.9 us with 2, the halo rows
// being a smaller share), 1 in double precision (register footprint of the
// paired loads). Independent of the node tiles' R.
inline int pRows(bool f32, int rows) {
    (void)rows;
    return f32 ? 4 : 1;
}
// Adjacent columns per lane of the p' list's tiles (one vector access per
// array and row): 4 in single precision, 16-byte accesses, so a wave's row
// is 1 KB of contiguous memory (measured on 8 x 1024x512 slots rotating
// over 3 sets: the PCG stencil 26.2 against 28.8 us with 2 columns per
// lane, 56 against 51 % of 8 TB/s); 2 in double precision. A tile is 64 Cp
// columns wide.
constexpr int pColsOf(bool f32) { return f32 ? 4 : 2; }
inline int pTileW(bool f32) { return kTileW * pColsOf(f32); }
hipError_t launchPlane(int nreq, hipStream_t s, const SlotDesc* desc, const PlaneReq* req, PlaneOut* out);

// Shared by the library's translation units (defined in cp_engine.hip):
// record the message cp_last_error() returns and pass `code` through; the
// check that the current device is a gfx950.
int reportError(int code, const char* fmt, ...);
bool currentDeviceIsGfx950();

}  // namespace cpdev
