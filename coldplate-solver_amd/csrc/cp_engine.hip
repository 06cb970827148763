// cp_engine.hip — host engine and C ABI of the gfx950 SIMPLE backend.
//
// A Batch owns one HBM arena, a slot table (SlotDesc / SlotState arrays on
// the device), the tile -> slot list every kernel iterates over, and one HIP
// stream. A single-geometry context (cp_ctx) is a Batch with one slot.
//
// Two ways to drive a step:
//  * polled: the drop-in mode. The data-dependent inner loops (Jacobi sweeps,
//    PCG iterations) are launched in growing chunks and the host reads the
//    slot flags between chunks, so the reference's stop tests and caps are
//    reproduced without launching the full 200 + 2x20 kernels per step.
//  * graph: the batched / served mode. One outer iteration with fixed inner
//    budgets (sweeps, PCG iterations, refinement cycles) is captured once as
//    a hipGraph and replayed; slots whose inner loop already stopped, or
//    which have converged, are skipped on the device, so no host
//    synchronisation happens inside a replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/coldplate_defaults.h"
#include "../../include/coldplate_hip.h"
#include "cp_device.h"


using namespace cpdev;

namespace {

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
std::mutex g_errMu;
std::string g_lastError;

int setError(std::string* dst, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    {
        std::lock_guard<std::mutex> lk(g_errMu);
        g_lastError = buf;
    }
    if (dst) *dst = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return setError(&err, CP_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

size_t roundUp(size_t x, size_t a) { return (x + a - 1) / a * a; }

// The kernels are built for gfx950 (MI355X) only: a device of any other
// architecture cannot run them.
bool isGfx950(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
    return strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

}  // namespace

int cpdev::reportError(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return setError(nullptr, code, "%s", buf);
}

bool cpdev::currentDeviceIsGfx950() {
    int dev = 0;
    return hipGetDevice(&dev) == hipSuccess && isGfx950(dev);
}

namespace {

// ---------------------------------------------------------------------------
// arena: first-fit free list over one device allocation (offsets in bytes,
// 256-byte granules), coalescing on free.
// ---------------------------------------------------------------------------
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> freeList;  // offset -> length

    void init(char* b, size_t s) {
        base = b;
        size = s;
        freeList.clear();
        freeList[0] = s;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = roundUp(bytes, 256);
        for (auto it = freeList.begin(); it != freeList.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t o2 = it->first + bytes;
                freeList.erase(it);
                if (rest) freeList[o2] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = roundUp(bytes, 256);
        auto it = freeList.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != freeList.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            freeList.erase(nx);
        }
        if (it != freeList.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                freeList.erase(it);
            }
        }
    }
    size_t largestFree() const {
        size_t m = 0;
        for (auto& kv : freeList) m = std::max(m, kv.second);
        return m;
    }
};

// Per-slot host bookkeeping.
struct SlotHost {
    bool used = false;
    cp_params prm{};
    size_t off = 0, bytes = 0;
    int M = 0, N = 0, pitch = 0;
    int tilesX = 0, tilesY = 0;
};

constexpr int kStateArrays = 12;  // u v p us vs pc dE dN aU aV gam bm
constexpr int kMgArrays = 6;      // aP cE cN b x z of each coarse multigrid level

// Shapes of the aggregation-multigrid hierarchy of an M x N cell grid: level
// 0 is the (M-1) x (N-2) p' grid, each next level aggregates 2x2 blocks, and
// coarsening stops once a level has at most kCoarseMax unknowns (at least one
// coarse level). Returns the number of levels.
// Over-relaxation of the SOR backend (pressure_solver.cpp:810-816): the
// caller's in (0, 2), else 2 / (1 + sqrt(1 - cos^2(pi / max(M, N))))
// clamped to [1, 1.95].
double sorOmegaFor(double w, int M, int N) {
    if (w > 0.0 && w < 2.0) return w;
    const double rs = std::cos(M_PI / double(std::max(M, N)));
    const double om = 2.0 / (1.0 + std::sqrt(1.0 - rs * rs));
    return std::max(1.0, std::min(1.95, om));
}

int mgShapes(int M, int N, int* H, int* W) {
    H[0] = M - 1;
    W[0] = N - 2;
    int n = 1;
    while ((n == 1 || long(H[n - 1]) * W[n - 1] > kCoarseMax) && n < kMaxLevels) {
        H[n] = (H[n - 1] + 1) / 2;
        W[n] = (W[n - 1] + 1) / 2;
        ++n;
    }
    return n;
}

size_t mgPitch(int W) { return roundUp(size_t(W) + 2, 32); }

// Device bytes of one slot: the node-grid arrays, then the coarse multigrid
// levels and the coarsest dense inverse (allocated for every slot, so that a
// footprint depends on the grid and the arithmetic mode alone).
size_t slotMainBytes(int precision, int M, int N) {
    const size_t tE = precision == CP_PREC_MIXED ? sizeof(float) : sizeof(double);
    const int presArrays = precision == CP_PREC_MIXED ? P_COUNT : P_COUNT - 1;  // FP64: x aliases p'
    const size_t perNode = kStateArrays * sizeof(double) + (2 * M_COUNT + presArrays) * tE;
    const size_t pitch = roundUp(size_t(N) + 1, 32);
    return roundUp(size_t(M + 1) * pitch * perNode + 256 * 32, 256);
}

size_t slotMgBytes(int precision, int M, int N) {
    const size_t tE = precision == CP_PREC_MIXED ? sizeof(float) : sizeof(double);
    int H[kMaxLevels], W[kMaxLevels];
    const int n = mgShapes(M, N, H, W);
    size_t bytes = roundUp(sizeof(double) * kCoarseMax * kCoarseMax, 256);
    bytes += 2 * roundUp(sizeof(double) * CP_DP_WINDOW_MAX, 256);  // pressure-drop window and slope buffer
    for (int l = 1; l < n; ++l) bytes += kMgArrays * roundUp(size_t(H[l] + 2) * mgPitch(W[l]) * tE, 256);
    return bytes;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------
struct cp_batch {
    std::string err;
    int precision = CP_PREC_FP64;
    int maxSlots = 0;
    int device = 0;
    int numCU = 256;
    hipStream_t stream = nullptr;
    Arena arena;
    char* arenaBase = nullptr;
    size_t maxTiles = 0;
    // device tables
    SlotDesc* dDesc = nullptr;
    SlotState* dState = nullptr;
    int* dTileSlot = nullptr;
    int* dNTiles = nullptr;
    int* dPTileSlot = nullptr;  // p' unknowns' tile list
    int* dPNTiles = nullptr;
    // multigrid tile lists: level l's list at dMgTileSlot + l * maxTiles
    int* dMgTileSlot = nullptr;
    int* dMgNTiles = nullptr;
    int mgLevels = 0;  // deepest hierarchy of any multigrid slot admitted so far
    int mgTail = 0;    // deepest first tail level of any multigrid slot admitted so far
    int dpAny = 0;     // some slot admitted so far uses the pressure-drop exit
    // backend families some slot admitted so far uses (sticky, like the
    // multigrid shape: the captured step graph keeps its launches)
    int anyExplicit = 0, anyImplicit = 0, anySOR = 0, anyCG = 0;
    // CG slots by preconditioner (k_pcg_q launches one form per kind)
    int anyMgCG = 0, anyJacCG = 0;
    double* dPartials = nullptr;
    PlaneReq* dPlaneReq = nullptr;
    PlaneOut* dPlaneOut = nullptr;
    // host mirrors
    std::vector<SlotDesc> hDesc;
    std::vector<SlotHost> slots;
    SlotState* hState = nullptr;  // pinned
    PlaneReq* hPlaneReq = nullptr;
    PlaneOut* hPlaneOut = nullptr;
    int totalTiles = 0;
    int totalPTiles = 0;
    // graph mode
    hipGraphExec_t graphExec = nullptr;
    int graphJ = -1, graphP = -1, graphR = -1, graphL = -1;
    // measurement: timing bursts inside cp_serve (every kProbeEvery polls)
    // of the kernels in probeMask (bit k: kernel k), one record per burst
    int probeMask = 0;
    std::vector<hipEvent_t> burstEv;  // event pair of the timing bursts
    struct ProbeRec {
        int kernel, launches;
        double ms, units;  // whole burst: elapsed time, work units over all launches
    };
    std::vector<ProbeRec> probes;
    // tile height R (rows per wave) shared by every slot: 1, 2, 4 or 8
    int tileRows = 4;
    // workgroups per CU of the replayed graph's launches
    int wgPerCU = 16;
    // rows per wave of the finest multigrid level's tiles (0: from tileRows)
    int mgTileRows = 0;
    // tile order of the launches (KArgs::order), a mask: bit 0 for the node
    // kernels, bit 1 for the p' kernels, bit 2 for the multigrid levels;
    // 1 = launch order
    int tileOrder = 0;
    int mgRowsEff() const {
        return mgTileRows > 0 ? mgTileRows : mgRows(precision == CP_PREC_MIXED, tileRows);
    }
    // replay the captured step graph (1) or launch its kernels directly (0)
    int useGraph = 1;
    // KArgs::fold of the launches being enqueued (set by the fixed-budget
    // step only)
    int fold = 0;
    // host mirrors of a batch laid out in the static device scratch (no
    // pinned allocations): hState / hPlaneReq / hPlaneOut point here
    SlotState staticState;
    PlaneReq staticPlaneReq[2];
    PlaneOut staticPlaneOut[2];

    size_t tElem() const { return precision == CP_PREC_MIXED ? sizeof(float) : sizeof(double); }
    size_t bytesPerNode() const {
        // state arrays (double) + momentum U/V (8 each) + pressure arrays
        const int presArrays = precision == CP_PREC_MIXED ? P_COUNT : P_COUNT - 1;  // FP64: x aliases p'
        return kStateArrays * sizeof(double) + (2 * M_COUNT + presArrays) * tElem();
    }
    size_t mainBytes(int M, int N) const { return slotMainBytes(precision, M, N); }
    size_t footprint(int M, int N) const { return slotMainBytes(precision, M, N) + slotMgBytes(precision, M, N); }
};

struct cp_ctx {
    cp_batch* b = nullptr;
    int slot = 0;
    bool ghostsDirty = false;
};

namespace {

// Workgroups of a launch: enough for the tiles, at most wgPerCU per CU, and
// a multiple of the XCD count (the kernels split the tile list per XCD).
int gridFor(const cp_batch* b, int tiles) {
    const int cap = b->numCU * b->wgPerCU;
    const int g = std::max(1, std::min(tiles, cap));
    return (g + kNumXCD - 1) / kNumXCD * kNumXCD;
}

KArgs kargs(cp_batch* b) {
    KArgs a;
    a.desc = b->dDesc;
    a.state = b->dState;
    a.tileSlot = b->dTileSlot;
    a.nTiles = b->dNTiles;
    a.partials = b->dPartials;
    a.force = 0;
    a.level = 0;
    a.color = 0;
    a.order = b->tileOrder & 1;
    a.fold = b->fold;
    a.pvar = kPcgQAll;
    return a;
}

// kernels that read and write p' unknowns only run over the p' tile list
bool pOnly(Kern k) { return k == K_PCG_Q || k == K_PCG_U || k == K_REF_ACC || k == K_REF_RES; }

KArgs kargsFor(cp_batch* b, Kern k) {
    KArgs a = kargs(b);
    if (pOnly(k)) {
        a.tileSlot = b->dPTileSlot;
        a.nTiles = b->dPNTiles;
        a.order = (b->tileOrder >> 1) & 1;
    }
    return a;
}

int launch(cp_batch* b, Kern k, int grid) {
    const bool f32 = b->precision == CP_PREC_MIXED;
    const int rows = pOnly(k) ? pRows(f32, b->tileRows) : b->tileRows;
    const hipError_t e = launchKernel(k, f32, rows, grid, b->maxSlots, b->stream, kargsFor(b, k));
    if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return CP_OK;
}

// The k_pcg_q launches of one CG iteration: in single precision one launch
// of one form (KArgs::pvar) per preconditioner kind the batch holds, at the
// first iteration of a CG (every running slot's pk is 0) or a later one;
// in double precision one launch for every slot. With `force` (timing
// bursts) the later-iteration forms, on the given stream.
// fold: the form that also adds the two updates before it to x (kPcgQFold)
int launchPcgQ(cp_batch* b, int grid, bool first, bool force = false, hipStream_t stream = nullptr, bool fold = false) {
    const bool f32 = b->precision == CP_PREC_MIXED;
    hipStream_t s = stream ? stream : b->stream;
    KArgs a = kargsFor(b, K_PCG_Q);
    a.force = force ? 1 : 0;
    const int rows = pRows(f32, b->tileRows);
    const int vars[2] = {b->anyMgCG ? kPcgQMg : -1, b->anyJacCG ? kPcgQJac : -1};
    for (int v : vars) {
        if (f32 && v < 0) continue;
        a.pvar = f32 ? (v | (first ? kPcgQFirst : 0) | (fold ? kPcgQFold : 0)) : kPcgQAll;
        const hipError_t e = launchWorkOnly(K_PCG_Q, f32, rows, grid, s, a);
        if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
        if (!f32) break;
    }
    return CP_OK;
}

// ---- multigrid launches ---------------------------------------------------
// Workgroups of a level's launch: the full grid on every level. Coarse-level
// tiles are short dependency chains (load, two sweeps, restriction), so
// they need many workgroups in flight; surplus workgroups exit at once.
int mgGrid(const cp_batch* b, int level) {
    (void)level;
    return gridFor(b, 1 << 30);
}

// finalPhase >= 0: launch the work kernel alone, then the finalisation of
// that phase instead of the kernel's own.
int launchMg(cp_batch* b, Kern k, int level, int finalPhase = -1) {
    const bool f32 = b->precision == CP_PREC_MIXED;
    const bool perSlot = k == K_MG_FACTOR || k == K_MG_TAIL;
    KArgs a = kargs(b);
    a.tileSlot = b->dMgTileSlot + size_t(level) * b->maxTiles;
    a.nTiles = b->dMgNTiles + level;
    a.level = level;
    a.order = (b->tileOrder >> 2) & 1;
    const int rows = level == 0 ? b->mgRowsEff() : mgCoarseRows(f32);
    const int grid = perSlot ? b->maxSlots : mgGrid(b, level);
    hipError_t e;
    if (finalPhase < 0) {
        e = launchKernel(k, f32, rows, grid, b->maxSlots, b->stream, a);
    } else {
        e = launchWorkOnly(k, f32, rows, grid, b->stream, a);
        if (e == hipSuccess) e = launchFinalize(finalPhase, b->maxSlots, b->stream, kargs(b));
    }
    if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return CP_OK;
}

// Coarse operators of every level and the coarsest inverse (after assembly).
int mgSetup(cp_batch* b) {
    int rc;
    // grid-wide launches down to the batch's first tail level; the levels
    // below it are coarsened in LDS by the per-slot factor launch
    const int last = std::min(b->mgTail, b->mgLevels - 1);
    for (int l = 1; l <= last; ++l)
        if ((rc = launchMg(b, K_MG_COARSEN, l))) return rc;
    return b->mgLevels > 0 ? launchMg(b, K_MG_FACTOR, last + 1) : CP_OK;
}

// z = V-cycle(r) for every multigrid slot whose CG is running, then r.z:
// per-level launches above the tails, one tail launch, per-level launches
// back up.
// afterUpdate: the V-cycle follows a PCG update launched without its
// finalisation; the finest up-sweep is then finalised by PH_PCGU_MGZ, which
// applies the update's stop test first (one launch fewer per iteration).
int mgVcycle(cp_batch* b, bool afterUpdate = false) {
    if (b->mgLevels == 0) return CP_OK;
    const int T0 = b->mgTail;
    int rc;
    for (int l = 0; l < T0; ++l)
        if ((rc = launchMg(b, K_MG_DOWN, l))) return rc;
    if ((rc = launchMg(b, K_MG_TAIL, 0))) return rc;
    for (int l = T0 - 1; l >= 0; --l)
        if ((rc = launchMg(b, K_MG_UP, l, (l == 0 && afterUpdate) ? int(PH_PCGU_MGZ) : -1))) return rc;
    return CP_OK;
}

int pollState(cp_batch* b) {
    std::string& err = b->err;
    HIP_TRY(hipMemcpyAsync(b->hState, b->dState, sizeof(SlotState) * b->maxSlots, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return CP_OK;
}

bool anyRunning(const cp_batch* b) {
    for (int s = 0; s < b->maxSlots; ++s)
        if (b->slots[s].used && b->hState[s].status == ST_RUNNING) return true;
    return false;
}

// Rebuild the tile -> slot list and upload descriptors.
int rebuildTiles(cp_batch* b) {
    std::string& err = b->err;
    std::vector<int> ts;
    for (int s = 0; s < b->maxSlots; ++s) {
        if (!b->slots[s].used) continue;
        SlotDesc& d = b->hDesc[s];
        d.tile0 = int(ts.size());
        for (int k = 0; k < d.ntiles; ++k) ts.push_back(s);
    }
    if (ts.size() > b->maxTiles) return setError(&err, CP_E_CAPACITY, "tile table full");
    b->totalTiles = int(ts.size());
    std::vector<int> pts;
    for (int s = 0; s < b->maxSlots; ++s) {
        if (!b->slots[s].used) continue;
        SlotDesc& d = b->hDesc[s];
        d.ptile0 = int(pts.size());
        for (int k = 0; k < d.pntiles; ++k) pts.push_back(s);
    }
    if (pts.size() > b->maxTiles) return setError(&err, CP_E_CAPACITY, "tile table full");
    b->totalPTiles = int(pts.size());
    for (int s = 0; s < b->maxSlots; ++s)
        if (b->slots[s].used && (b->hDesc[s].dpConv || b->hDesc[s].slopeGate)) b->dpAny = 1;
    for (int s = 0; s < b->maxSlots; ++s) {
        if (!b->slots[s].used) continue;
        const SlotDesc& d = b->hDesc[s];
        (d.momType == 0 ? b->anyExplicit : b->anyImplicit) = 1;
        (d.presType == 0 ? b->anySOR : b->anyCG) = 1;
        if (d.presType != 0) (d.precond ? b->anyMgCG : b->anyJacCG) = 1;
    }
    int counts[kMaxLevels];
    std::vector<std::vector<int>> lists(kMaxLevels);  // alive until the stream is synchronised
    // mgLevels never decreases: the captured step graph keeps its shape as
    // slots come and go (levels deeper than every admitted slot's hierarchy
    // have empty tile lists, and their launches exit at once)
    for (int l = 0; l < kMaxLevels; ++l) {
        std::vector<int>& ml = lists[l];
        for (int s = 0; s < b->maxSlots; ++s) {
            if (!b->slots[s].used) continue;
            SlotDesc& d = b->hDesc[s];
            if (l >= d.nLevels) continue;
            d.lv[l].tile0 = int(ml.size());
            for (int k = 0; k < d.lv[l].ntiles; ++k) ml.push_back(s);
            b->mgLevels = std::max(b->mgLevels, d.nLevels);
            b->mgTail = std::max(b->mgTail, d.mgTail);
        }
        if (ml.size() > b->maxTiles) return setError(&err, CP_E_CAPACITY, "multigrid tile table full");
        counts[l] = int(ml.size());
        if (!ml.empty())
            HIP_TRY(hipMemcpyAsync(b->dMgTileSlot + size_t(l) * b->maxTiles, ml.data(), ml.size() * sizeof(int),
                                   hipMemcpyHostToDevice, b->stream));
    }
    HIP_TRY(hipMemcpyAsync(b->dMgNTiles, counts, sizeof(counts), hipMemcpyHostToDevice, b->stream));
    if (!ts.empty())
        HIP_TRY(hipMemcpyAsync(b->dTileSlot, ts.data(), ts.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->dNTiles, &b->totalTiles, sizeof(int), hipMemcpyHostToDevice, b->stream));
    if (!pts.empty())
        HIP_TRY(hipMemcpyAsync(b->dPTileSlot, pts.data(), pts.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->dPNTiles, &b->totalPTiles, sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->dDesc, b->hDesc.data(), sizeof(SlotDesc) * b->maxSlots, hipMemcpyHostToDevice,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return CP_OK;
}

void defaultInlet(const SlotDesc& d, SlotState& st) {
    st.inletVelocity = (d.inletRamp && d.rampSteps > 0 && st.iter < d.rampSteps)
                           ? (double(st.iter + 1) / double(d.rampSteps)) * d.targetVel
                           : d.targetVel;
}

// Host-side geometry preparation (SIMPLE.cpp:556-630, SIMPLE.h:546-590):
// gamma = 1 - cellType; alpha = alphaMax (1-g)/(1+q g+1e-10); rows 0 and M-2
// forced solid; alpha averaged onto the u and v faces.
void buildGeometry(const cp_params& p, const double* ct, int pitch, double alphaMax, std::vector<double>& aU,
                   std::vector<double>& aV, std::vector<double>& gam, std::vector<double>* gammaOut,
                   std::vector<double>* alphaOut) {
    const int M = p.M, N = p.N;
    std::vector<double> g(size_t(M) * N), al(size_t(M) * N);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) {
            const double gv = 1.0 - ct[size_t(i) * N + j];
            g[size_t(i) * N + j] = gv;
            al[size_t(i) * N + j] = alphaMax * (1.0 - gv) / (1.0 + p.brinkmanQ * gv + 1e-10);
        }
    if (M >= 2)
        for (int j = 0; j < N; ++j) {
            al[j] = alphaMax;
            al[size_t(M - 2) * N + j] = alphaMax;
            g[j] = 0.0;
            g[size_t(M - 2) * N + j] = 0.0;
        }
    const size_t plane = size_t(M + 1) * pitch;
    aU.assign(plane, 0.0);
    aV.assign(plane, 0.0);
    gam.assign(plane, 0.0);
    for (int i = 1; i <= M - 1; ++i)
        for (int j = 1; j <= N - 2; ++j)
            aU[size_t(i) * pitch + j] = 0.5 * (al[size_t(i - 1) * N + j - 1] + al[size_t(i - 1) * N + j]);
    for (int i = 1; i <= M - 2; ++i)
        for (int j = 1; j <= N - 1; ++j)
            aV[size_t(i) * pitch + j] = 0.5 * (al[size_t(i - 1) * N + j - 1] + al[size_t(i) * N + j - 1]);
    for (int i = 1; i <= M; ++i)
        for (int j = 1; j <= N; ++j) gam[size_t(i) * pitch + j] = g[size_t(i - 1) * N + j - 1];
    if (gammaOut) *gammaOut = g;
    if (alphaOut) *alphaOut = al;
}

int validate(const cp_params* p, std::string* err) {
    if (!p) return setError(err, CP_E_ARG, "null params");
    if (p->M < 4 || p->N < 4) return setError(err, CP_E_ARG, "grid must be at least 4 x 4 cells (M=%d N=%d)", p->M, p->N);
    if (!(p->hx > 0) || !(p->hy > 0)) return setError(err, CP_E_ARG, "hx, hy must be > 0");
    if (!(p->rho > 0) || !(p->mu > 0)) return setError(err, CP_E_ARG, "rho, mu must be > 0");
    if (!(p->uvAlpha > 0) || !(p->uvAlpha <= 1)) return setError(err, CP_E_ARG, "uvAlpha must be in (0, 1]");
    if (p->precision != CP_PREC_FP64 && p->precision != CP_PREC_MIXED)
        return setError(err, CP_E_ARG, "unknown precision %d", p->precision);
    if (p->pressurePrecond != CP_PRECOND_JACOBI && p->pressurePrecond != CP_PRECOND_AMG)
        return setError(err, CP_E_ARG, "unknown pressurePrecond %d", p->pressurePrecond);
    if (p->pressureSolverType < 0 || p->pressureSolverType > 4)
        return setError(err, CP_E_ARG, "unknown pressureSolverType %d (0 = SOR, 1-4 = CG)", p->pressureSolverType);
    if (p->momentumSolverType != 0 && p->momentumSolverType != 1)
        return setError(err, CP_E_ARG, "unknown momentumSolverType %d (0 = explicit, 1 = implicit)",
                        p->momentumSolverType);
    if (p->pressurePrecond == CP_PRECOND_AMG && p->pressureSolverType != 0 &&
        (!(p->mgOmega > 0) || !(p->mgOmega <= 1) || !(p->mgScale > 0) || !(p->mgScale <= 4)))
        return setError(err, CP_E_ARG, "mgOmega must be in (0, 1] and mgScale in (0, 4]");
    if (p->usePressureDropConvergence && !(p->dpConvergenceWindow >= 1 && p->dpConvergenceWindow <= CP_DP_WINDOW_MAX))
        return setError(err, CP_E_ARG, "dpConvergenceWindow must be in [1, %d]", CP_DP_WINDOW_MAX);
    if (p->usePressureDropSlopeGate && !(p->dpSlopeWindowIters >= 0 && p->dpSlopeWindowIters <= CP_DP_WINDOW_MAX))
        return setError(err, CP_E_ARG, "dpSlopeWindowIters must be in [0, %d]", CP_DP_WINDOW_MAX);
    if (p->pressurePrecond == CP_PRECOND_AMG && p->pressureSolverType != 0) {
        int H[kMaxLevels], W[kMaxLevels];
        const int n = mgShapes(p->M, p->N, H, W);
        if (long(H[n - 1]) * W[n - 1] > kCoarseMax)
            return setError(err, CP_E_ARG, "grid too elongated for %d multigrid levels", kMaxLevels);
    }
    return CP_OK;
}

// Rows per wave for a grid: taller tiles cut the stencil kernels' vertical
// re-reads; small grids keep short tiles so the launch still has enough
// workgroups. An explicit choice (the autotuner's) overrides this.
int defaultTileRows(int M, int N) {
    const long cells = long(M) * long(N);
    if (cells >= (1L << 18)) return 4;
    if (cells >= (1L << 15)) return 2;
    return 1;
}

void setTiling(SlotDesc& d, int tileRows, int mgRows0) {
    d.tileRows = tileRows;
    d.tilesX = (d.N + 1 + kTileW - 1) / kTileW;
    const int rowsPerTile = kTileH * d.tileRows;
    const int tilesY = (d.M + 1 + rowsPerTile - 1) / rowsPerTile;
    d.ntiles = d.tilesX * tilesY;
    // p' unknowns: rows 1 .. M-1 from row 1, columns 1 .. N-2 from column 0
    const int pRowsPerTile = kTileH * pRows(d.mixed != 0, d.tileRows);
    d.ptilesX = (d.N - 1 + pTileW(d.mixed != 0) - 1) / pTileW(d.mixed != 0);
    d.pntiles = d.ptilesX * ((d.M - 1 + pRowsPerTile - 1) / pRowsPerTile);
    for (int l = 0; l < d.nLevels; ++l) {
        MgLevel& L = d.lv[l];
        const int rows = l == 0 ? mgRows0 : mgCoarseRows(d.mixed != 0);
        L.tilesX = (L.W + kMgTileCols - 1) / kMgTileCols;
        L.ntiles = L.tilesX * ((L.H + rows - 1) / rows);
    }
}

SlotDesc makeDesc(const cp_params& p, int precision, int tileRows, int mgRows0) {
    SlotDesc d;
    memset(&d, 0, sizeof(d));
    d.M = p.M;
    d.N = p.N;
    d.pitch = int(roundUp(size_t(p.N) + 1, 32));
    d.mixed = precision == CP_PREC_MIXED ? 1 : 0;
    d.dpConv = p.usePressureDropConvergence ? 1 : 0;
    d.dpWindow = p.dpConvergenceWindow;
    d.slopeGate = p.usePressureDropSlopeGate ? 1 : 0;
    d.slopeWindow = p.dpSlopeWindowIters;
    d.dpPercent = p.dpConvergencePercent;
    d.slopeMaxDelta = p.dpSlopeMaxDelta;
    d.xCoreIn = p.xPlaneCoreInlet;
    d.xCoreOut = p.xPlaneCoreOutlet;
    d.presType = p.pressureSolverType == 0 ? 0 : 1;
    d.momType = p.momentumSolverType == 0 ? 0 : 1;
    d.dE32 = (precision == CP_PREC_MIXED && d.momType == 1) ? 1 : 0;
    d.sorOmega = sorOmegaFor(p.sorOmega, p.M, p.N);
    d.precond = d.presType == 0 ? CP_PRECOND_JACOBI : p.pressurePrecond;  // SOR slots build no hierarchy
    d.mgOmega = p.mgOmega;
    d.mgScale = p.mgScale;
    if (d.precond == CP_PRECOND_AMG) {
        int H[kMaxLevels], W[kMaxLevels];
        d.nLevels = mgShapes(p.M, p.N, H, W);
        for (int l = 0; l < d.nLevels; ++l) {
            d.lv[l].H = H[l];
            d.lv[l].W = W[l];
            d.lv[l].pitch = l == 0 ? d.pitch : int(mgPitch(W[l]));
        }
        // the tail kernel takes over from the first level whose remaining
        // levels fit its LDS arrays (the coarsest alone always does)
        const long cap = precision == CP_PREC_MIXED ? kTailCapF32 : kTailCapF64;
        d.mgTail = d.nLevels - 1;
        for (int l = 1; l < d.nLevels; ++l) {
            long sum = 0;
            for (int k = l; k < d.nLevels; ++k) sum += long(H[k]) * W[k];
            if (sum <= cap) {
                d.mgTail = l;
                break;
            }
        }
    }
    setTiling(d, tileRows, mgRows0);
    d.nU = (p.M - 1) * (p.N - 2);
    d.nV = (p.M - 2) * (p.N - 1);
    d.convScheme = p.convectionScheme;
    d.pseudo = p.enablePseudoTime;
    d.localPseudo = p.useLocalPseudoTime;
    d.maxMomIter = p.fixedMomentumSweeps > 0 ? p.fixedMomentumSweeps : p.maxMomentumIter;
    d.maxPresIter = p.maxPressureIter;
    d.mixed = precision == CP_PREC_MIXED ? 1 : 0;
    d.refineCycles = d.mixed ? std::max(1, p.refineCycles) : 1;
    d.massScope = p.massResidualScope;
    d.maxIterations = p.maxIterations;
    d.inletRamp = p.enableInletRamp;
    d.rampSteps = p.rampSteps;
    d.enableSER = p.enableSER;
    d.serMinIter = p.serMinIter;
    d.serUseMax = p.serUseMaxResid;
    d.enableLS = p.enableLineSearch;
    d.lsMaxTries = p.lsMaxTries;
    d.enableRamp = p.enableCflRamp;
    d.hx = p.hx;
    d.hy = p.hy;
    d.rho = p.rho;
    d.vol = p.hx * p.hy;
    d.De = p.mu * p.hy / p.hx;
    d.Dn = p.mu * p.hx / p.hy;
    d.convScale = p.enableConvectionScaling ? (6.0 / 7.0) : 1.0;
    double Ht = p.Ht;
    if (p.enableTwoPointFiveD && Ht <= 0.0) Ht = p.hy;  // SIMPLE.cpp:55-60
    const double sinkCoeff =
        (p.enableTwoPointFiveD && Ht > 0.0) ? p.sinkMultiplier * (5.0 * p.mu / (2.0 * Ht * Ht)) : 0.0;
    d.sinkDiag = sinkCoeff * d.vol;
    d.uvAlpha = p.uvAlpha;
    d.pAlpha = p.pAlpha;
    d.hChar = std::min(p.hx, p.hy);
    d.timeStep = p.timeStepMultiplier * (double(p.N) * p.hx) / std::max(p.U_in, 1e-6);  // SIMPLE.cpp:68-70
    d.minSpeed = p.minPseudoSpeed;
    d.momTol = p.fixedMomentumSweeps > 0 ? 0.0 : p.momentumTol;
    d.pTol = p.pTol;
    d.epsilon = p.epsilon;
    d.targetVel = p.U_in;
    d.serCFLMin = p.serCFLMin;
    d.serCFLMax = p.serCFLMax;
    d.serSmooth = p.serSmooth;
    d.serMaxInc = p.serMaxIncrease;
    d.serMaxDec = p.serMaxDecrease;
    d.lsAlphaMin = p.lsAlphaMin;
    d.lsAlphaReduce = p.lsAlphaReduce;
    d.lsTol = p.lsResidIncreaseTol;
    d.cflInit = p.pseudoCFLInitial;
    d.cflMax = p.pseudoCFLMax;
    d.rampStart = p.cflRampStartRes;
    d.rampExp = p.cflRampExponent;
    d.rampSmooth = p.cflRampSmooth;
    return d;
}

double alphaMaxOf(const cp_params& p) {
    return p.alphaMax > 0 ? p.alphaMax : p.mu / std::max(p.KMin, 1e-20);  // SIMPLE.cpp:65-66
}

void assignPointers(cp_batch* b, SlotDesc& d, char* base) {
    const size_t plane = size_t(d.M + 1) * d.pitch;
    double* sp = reinterpret_cast<double*>(base);
    double** st[kStateArrays] = {&d.u, &d.v, &d.p, &d.us, &d.vs, &d.pc, &d.dE, &d.dN, &d.aU, &d.aV, &d.gam, &d.bm};
    for (int k = 0; k < kStateArrays; ++k) *st[k] = sp + k * plane;
    char* tp = base + kStateArrays * plane * sizeof(double);
    const size_t tb = plane * b->tElem();
    for (int k = 0; k < M_COUNT; ++k) d.mu[k] = tp + k * tb;
    for (int k = 0; k < M_COUNT; ++k) d.mv[k] = tp + (M_COUNT + k) * tb;
    const int presArrays = b->precision == CP_PREC_MIXED ? P_COUNT : P_COUNT - 1;
    for (int k = 0; k < presArrays; ++k) d.pr[k] = tp + (2 * M_COUNT + k) * tb;
    if (b->precision != CP_PREC_MIXED) d.pr[P_X] = d.pc;  // FP64: PCG writes p' directly
    // multigrid: the finest level is the p' system itself; coarse levels and
    // the dense inverse follow the slot's main arrays
    char* mp = base + b->mainBytes(d.M, d.N);
    d.coarseInv = reinterpret_cast<double*>(mp);
    mp += roundUp(sizeof(double) * kCoarseMax * kCoarseMax, 256);
    d.dpHist = reinterpret_cast<double*>(mp);
    mp += roundUp(sizeof(double) * CP_DP_WINDOW_MAX, 256);
    d.dpSlope = reinterpret_cast<double*>(mp);
    mp += roundUp(sizeof(double) * CP_DP_WINDOW_MAX, 256);
    if (d.nLevels > 0) {
        MgLevel& L0 = d.lv[0];
        L0.aP = d.pr[P_AP];
        L0.cE = d.pr[P_CE];
        L0.cN = d.pr[P_CN];
        L0.b = d.pr[P_R];
        L0.x = d.pr[P_MGX];
        L0.z = d.pr[P_Z];
        for (int l = 1; l < d.nLevels; ++l) {
            MgLevel& L = d.lv[l];
            const size_t arr = roundUp(size_t(L.H + 2) * L.pitch * b->tElem(), 256);
            void** ptrs[kMgArrays] = {&L.aP, &L.cE, &L.cN, &L.b, &L.x, &L.z};
            for (int k = 0; k < kMgArrays; ++k) {
                *ptrs[k] = mp;
                mp += arr;
            }
        }
    }
}

// Copy a pitched device plane to a dense rows x cols host array.
int downloadPlane(cp_batch* b, const double* dsrc, int pitch, int rows, int cols, double* out) {
    std::string& err = b->err;
    HIP_TRY(hipMemcpy2DAsync(out, sizeof(double) * cols, dsrc, sizeof(double) * pitch, sizeof(double) * cols, rows,
                             hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return CP_OK;
}

// the same for an array that holds floats (dE, dN of SlotDesc::dE32)
int downloadPlaneF(cp_batch* b, const double* dsrc, int pitch, int rows, int cols, double* out) {
    std::string& err = b->err;
    std::vector<float> tmp(size_t(rows) * cols);
    HIP_TRY(hipMemcpy2DAsync(tmp.data(), sizeof(float) * cols, dsrc, sizeof(float) * pitch, sizeof(float) * cols, rows,
                             hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (size_t q = 0; q < tmp.size(); ++q) out[q] = tmp[q];
    return CP_OK;
}

int uploadPlane(cp_batch* b, double* ddst, int pitch, int rows, int cols, const double* in) {
    std::string& err = b->err;
    HIP_TRY(hipMemcpy2DAsync(ddst, sizeof(double) * pitch, in, sizeof(double) * cols, sizeof(double) * cols, rows,
                             hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return CP_OK;
}

int batchCreate(int precision, size_t budget, int maxSlots, cp_batch** out, std::string* errOut) {
    if (!out) return setError(errOut, CP_E_ARG, "null output pointer");
    if (maxSlots < 1 || maxSlots > 4096) return setError(errOut, CP_E_ARG, "max_slots out of range");
    if (precision != CP_PREC_FP64 && precision != CP_PREC_MIXED)
        return setError(errOut, CP_E_ARG, "unknown precision %d", precision);
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return setError(errOut, CP_E_NODEV, "no HIP device");
    if (hipGetDevice(&dev) != hipSuccess || !isGfx950(dev))
        return setError(errOut, CP_E_NODEV, "the current HIP device %d is not a gfx950 (MI355X) device", dev);
    cp_batch* b = new cp_batch();
    std::string& err = b->err;
    b->precision = precision;
    b->maxSlots = maxSlots;
    HIP_TRY(hipGetDevice(&b->device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, b->device));
    b->numCU = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    budget = roundUp(std::max<size_t>(budget, 1 << 20), 256);
    HIP_TRY(hipMalloc(&b->arenaBase, budget));
    b->arena.init(b->arenaBase, budget);
    // tile capacity: every node costs bytesPerNode; a tile holds 256 nodes,
    // plus up to one partial tile row/column per slot.
    b->maxTiles = budget / (b->bytesPerNode() * kBlock / 4) + size_t(maxSlots) * 64 + 64;
    HIP_TRY(hipMalloc(&b->dDesc, sizeof(SlotDesc) * maxSlots));
    HIP_TRY(hipMalloc(&b->dState, sizeof(SlotState) * maxSlots));
    HIP_TRY(hipMalloc(&b->dTileSlot, sizeof(int) * b->maxTiles));
    HIP_TRY(hipMalloc(&b->dNTiles, sizeof(int)));
    HIP_TRY(hipMalloc(&b->dPTileSlot, sizeof(int) * b->maxTiles));
    HIP_TRY(hipMalloc(&b->dPNTiles, sizeof(int)));
    HIP_TRY(hipMalloc(&b->dMgTileSlot, sizeof(int) * b->maxTiles * kMaxLevels));
    HIP_TRY(hipMalloc(&b->dMgNTiles, sizeof(int) * kMaxLevels));
    HIP_TRY(hipMalloc(&b->dPartials, sizeof(double) * kNumPartials * b->maxTiles));
    HIP_TRY(hipMalloc(&b->dPlaneReq, sizeof(PlaneReq) * 2 * maxSlots));
    HIP_TRY(hipMalloc(&b->dPlaneOut, sizeof(PlaneOut) * 2 * maxSlots));
    HIP_TRY(hipHostMalloc(&b->hState, sizeof(SlotState) * maxSlots, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&b->hPlaneReq, sizeof(PlaneReq) * 2 * maxSlots, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&b->hPlaneOut, sizeof(PlaneOut) * 2 * maxSlots, hipHostMallocDefault));
    memset(b->hState, 0, sizeof(SlotState) * maxSlots);
    b->hDesc.assign(maxSlots, SlotDesc());
    memset(b->hDesc.data(), 0, sizeof(SlotDesc) * maxSlots);
    b->slots.assign(maxSlots, SlotHost());
    HIP_TRY(hipMemcpyAsync(b->dState, b->hState, sizeof(SlotState) * maxSlots, hipMemcpyHostToDevice, b->stream));
    const int zero = 0;
    HIP_TRY(hipMemcpyAsync(b->dNTiles, &zero, sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->dPNTiles, &zero, sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemsetAsync(b->dMgNTiles, 0, sizeof(int) * kMaxLevels, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    *out = b;
    return CP_OK;
}

void batchDestroy(cp_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graphExec) (void)hipGraphExecDestroy(b->graphExec);
    for (hipEvent_t ev : b->burstEv)
        if (ev) (void)hipEventDestroy(ev);
    (void)hipFree(b->arenaBase);
    (void)hipFree(b->dDesc);
    (void)hipFree(b->dState);
    (void)hipFree(b->dTileSlot);
    (void)hipFree(b->dNTiles);
    (void)hipFree(b->dPTileSlot);
    (void)hipFree(b->dPNTiles);
    (void)hipFree(b->dMgTileSlot);
    (void)hipFree(b->dMgNTiles);
    (void)hipFree(b->dPartials);
    (void)hipFree(b->dPlaneReq);
    (void)hipFree(b->dPlaneOut);
    (void)hipHostFree(b->hState);
    (void)hipHostFree(b->hPlaneReq);
    (void)hipHostFree(b->hPlaneOut);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int batchAdmit(cp_batch* b, const cp_params* prm, const double* ct, int* slotOut) {
    std::string& err = b->err;
    int rc = validate(prm, &err);
    if (rc) return rc;
    if (!ct || !slotOut) return setError(&err, CP_E_ARG, "null cellType or slot pointer");
    if (prm->precision != b->precision)
        return setError(&err, CP_E_ARG, "params precision %d differs from batch precision %d", prm->precision,
                        b->precision);
    int s = -1;
    for (int k = 0; k < b->maxSlots; ++k)
        if (!b->slots[k].used) { s = k; break; }
    if (s < 0) return setError(&err, CP_E_CAPACITY, "no free slot");
    SlotDesc d = makeDesc(*prm, b->precision, b->tileRows, b->mgRowsEff());
    if (b->totalTiles + size_t(d.ntiles) > b->maxTiles) return setError(&err, CP_E_CAPACITY, "tile table full");
    const size_t bytes = b->footprint(prm->M, prm->N);
    size_t off = 0;
    if (!b->arena.alloc(bytes, &off)) return setError(&err, CP_E_CAPACITY, "arena full (%zu bytes needed)", bytes);
    char* base = b->arenaBase + off;
    assignPointers(b, d, base);
    // zero the whole region: non-DOF entries of every array stay 0 for good
    hipError_t e = hipMemsetAsync(base, 0, bytes, b->stream);
    if (e != hipSuccess) {
        b->arena.release(off, bytes);
        return setError(&err, CP_E_HIP, "hipMemsetAsync: %s", hipGetErrorString(e));
    }
    std::vector<double> aU, aV, gam;
    buildGeometry(*prm, ct, d.pitch, alphaMaxOf(*prm), aU, aV, gam, nullptr, nullptr);
    const size_t plane = size_t(prm->M + 1) * d.pitch;
    std::vector<float> aU32, aV32;
    if (d.dE32) {
        // mixed-precision implicit slots read the Brinkman faces as floats
        aU32.assign(aU.begin(), aU.end());
        aV32.assign(aV.begin(), aV.end());
        HIP_TRY(hipMemcpyAsync(d.aU, aU32.data(), plane * sizeof(float), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(d.aV, aV32.data(), plane * sizeof(float), hipMemcpyHostToDevice, b->stream));
    } else {
        HIP_TRY(hipMemcpyAsync(d.aU, aU.data(), plane * sizeof(double), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(d.aV, aV.data(), plane * sizeof(double), hipMemcpyHostToDevice, b->stream));
    }
    HIP_TRY(hipMemcpyAsync(d.gam, gam.data(), plane * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    SlotHost& sh = b->slots[s];
    sh.used = true;
    sh.prm = *prm;
    sh.off = off;
    sh.bytes = bytes;
    sh.M = prm->M;
    sh.N = prm->N;
    sh.pitch = d.pitch;
    b->hDesc[s] = d;
    SlotState st;
    memset(&st, 0, sizeof(st));
    st.status = prm->maxIterations > 0 ? ST_RUNNING : ST_MAXITER;
    st.residU = st.residV = st.residMass = st.error = 1.0;  // SIMPLE.h:366-368
    st.pseudoCFL = prm->enableCflRamp ? prm->pseudoCFLInitial : 0.1;  // SIMPLE.cpp:79-82, SIMPLE.h:263
    st.lsAlpha = 1.0;
    defaultInlet(d, st);
    // state is written through the pinned mirror
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->hState[s] = st;
    HIP_TRY(hipMemcpyAsync(b->dState + s, &b->hState[s], sizeof(SlotState), hipMemcpyHostToDevice, b->stream));
    rc = rebuildTiles(b);
    if (rc) return rc;
    *slotOut = s;
    return CP_OK;
}

int batchEvict(cp_batch* b, int s) {
    std::string& err = b->err;
    if (s < 0 || s >= b->maxSlots || !b->slots[s].used) return setError(&err, CP_E_ARG, "slot %d not in use", s);
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->arena.release(b->slots[s].off, b->slots[s].bytes);
    b->slots[s] = SlotHost();
    memset(&b->hDesc[s], 0, sizeof(SlotDesc));
    memset(&b->hState[s], 0, sizeof(SlotState));
    HIP_TRY(hipMemcpyAsync(b->dState + s, &b->hState[s], sizeof(SlotState), hipMemcpyHostToDevice, b->stream));
    return rebuildTiles(b);
}

// ---- sub-steps, polled (drop-in semantics) --------------------------------
int maxOverSlots(cp_batch* b, int SlotDesc::*field) {
    int m = 0;
    for (int s = 0; s < b->maxSlots; ++s)
        if (b->slots[s].used) m = std::max(m, b->hDesc[s].*field);
    return m;
}

int polledMomentum(cp_batch* b) {
    const int grid = gridFor(b, b->totalTiles);
    int rc;
    if (b->anyExplicit) {
        if ((rc = launch(b, K_MOM_EXPL, grid)) || (rc = launch(b, K_MOM_EXPL_BC, grid))) return rc;
    }
    if (!b->anyImplicit) return CP_OK;
    if ((rc = launch(b, K_MOM_ASM, grid))) return rc;  // with the first sweep
    const int cap = maxOverSlots(b, &SlotDesc::maxMomIter);
    int launched = 1, chunk = 2;
    while (launched < cap) {
        const int n = std::min(chunk, cap - launched);
        for (int k = 0; k < n; ++k)
            if ((rc = launch(b, K_JAC, grid))) return rc;
        launched += n;
        if (launched >= cap) break;
        if ((rc = pollState(b))) return rc;
        bool pending = false;
        for (int s = 0; s < b->maxSlots; ++s)
            if (b->slots[s].used && b->hState[s].status == ST_RUNNING && !(b->hState[s].jdone[0] && b->hState[s].jdone[1]))
                pending = true;
        if (!pending) break;
        chunk = std::min(chunk * 2, 64);
    }
    return launch(b, K_MOM_FIN, grid);
}

// One SOR sweep pair (red, then black with the stop test).
int sorPair(cp_batch* b, int grid) {
    for (int color = 0; color < 2; ++color) {
        KArgs a = kargs(b);
        a.color = color;
        const hipError_t e =
            launchKernel(K_SOR, b->precision == CP_PREC_MIXED, b->tileRows, grid, b->maxSlots, b->stream, a);
        if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    }
    return CP_OK;
}

int polledSOR(cp_batch* b, int grid) {
    const int cap = maxOverSlots(b, &SlotDesc::maxPresIter);
    int launched = 0, chunk = 4, rc;
    while (launched < cap) {
        const int n = std::min(chunk, cap - launched);
        for (int k = 0; k < n; ++k)
            if ((rc = sorPair(b, grid))) return rc;
        launched += n;
        if (launched >= cap) break;
        if ((rc = pollState(b))) return rc;
        bool pending = false;
        for (int s = 0; s < b->maxSlots; ++s)
            if (b->slots[s].used && b->hState[s].status == ST_RUNNING && b->hDesc[s].presType == 0 &&
                !b->hState[s].sorDone)
                pending = true;
        if (!pending) break;
        chunk = std::min(chunk * 2, 64);
    }
    return CP_OK;
}

int polledPressure(cp_batch* b) {
    const int grid = gridFor(b, b->totalTiles);
    int rc = launch(b, K_PRES_ASM, grid);
    if (rc) return rc;
    if (b->anySOR && (rc = polledSOR(b, grid))) return rc;
    if (!b->anyCG) return CP_OK;
    if ((rc = mgSetup(b)) || (rc = mgVcycle(b))) return rc;
    const int cap = maxOverSlots(b, &SlotDesc::maxPresIter);
    const int cycles = maxOverSlots(b, &SlotDesc::refineCycles);
    for (int c = 0; c < cycles; ++c) {
        int launched = 0, chunk = 1;
        while (launched < cap) {
            const int n = std::min(chunk, cap - launched);
            for (int k = 0; k < n; ++k) {
                if ((rc = launchPcgQ(b, grid, launched + k == 0))) return rc;
                if ((rc = launch(b, K_PCG_U, grid))) return rc;
                if ((rc = mgVcycle(b))) return rc;
            }
            launched += n;
            if (launched >= cap) break;
            if ((rc = pollState(b))) return rc;
            bool pending = false;
            for (int s = 0; s < b->maxSlots; ++s)
                if (b->slots[s].used && b->hState[s].status == ST_RUNNING && !b->hState[s].pdone &&
                    !b->hState[s].stepPhase)
                    pending = true;
            if (!pending) break;
            chunk = std::min(chunk * 2, 16);
        }
        if (b->precision != CP_PREC_MIXED) break;
        if ((rc = launch(b, K_REF_ACC, grid))) return rc;
        if (c + 1 < cycles) {
            if ((rc = launch(b, K_REF_RES, grid))) return rc;
            if ((rc = mgVcycle(b))) return rc;
            if ((rc = pollState(b))) return rc;
            bool pending = false;
            for (int s = 0; s < b->maxSlots; ++s)
                if (b->slots[s].used && b->hState[s].status == ST_RUNNING && !b->hState[s].stepPhase) pending = true;
            if (!pending) break;
        }
    }
    return CP_OK;
}

// Pressure-drop window exit of every slot, after the controller.
int dpMonitor(cp_batch* b) {
    if (!b->dpAny) return CP_OK;
    const hipError_t e =
        launchKernel(K_DP_MONITOR, b->precision == CP_PREC_MIXED, 1, b->maxSlots, b->maxSlots, b->stream, kargs(b));
    if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return CP_OK;
}

int polledStep(cp_batch* b, bool outer) {
    int rc = polledMomentum(b);
    if (rc) return rc;
    if ((rc = polledPressure(b))) return rc;
    if ((rc = launch(b, outer ? K_CORR_OUTER : K_CORR, gridFor(b, b->totalTiles)))) return rc;
    return outer ? dpMonitor(b) : CP_OK;
}

// ---- graph mode (fixed inner budgets) -------------------------------------
// KArgs::fold of the fixed-budget step: a mixed-precision batch folds the
// momentum finish into the pressure assembly and the correction, and with
// one refinement cycle (p' = 0 + x) the accumulation into the correction.
// u*, v* and p' then hold no step's values; only the batch API runs this
// step, and it reads u, v and p alone (the context API, which reads every
// field, runs polled steps).
int stepFold(const cp_batch* b, int R) {
    if (b->precision != CP_PREC_MIXED) return 0;
    return (b->anyImplicit ? kFoldMom : 0) | (R == 1 ? kFoldPc : 0);
}

int enqueueFixedStepImpl(cp_batch* b, int J, int P, int R, int grid);

int enqueueFixedStep(cp_batch* b, int J, int P, int R, int grid) {
    b->fold = stepFold(b, R);
    const int rc = enqueueFixedStepImpl(b, J, P, R, grid);
    b->fold = 0;
    return rc;
}

int enqueueFixedStepImpl(cp_batch* b, int J, int P, int R, int grid) {
    int rc;
    if (b->anyExplicit) {
        if ((rc = launch(b, K_MOM_EXPL, grid)) || (rc = launch(b, K_MOM_EXPL_BC, grid))) return rc;
    }
    if (b->anyImplicit) {
        // the first sweep is the assembly's; the last sweep's stop test could
        // only end a system that the budget ends anyway, so it goes without
        // its finalisation and k_mom_finish takes the budget as the count
        const bool f32 = b->precision == CP_PREC_MIXED;
        for (int k = 0; k < J; ++k) {
            const Kern kk = k == 0 ? K_MOM_ASM : K_JAC;
            if (k + 1 < J) {
                if ((rc = launch(b, kk, grid))) return rc;
            } else {
                const hipError_t e = launchWorkOnly(kk, f32, b->tileRows, grid, b->stream, kargsFor(b, kk));
                if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
            }
        }
        if (J == 0 && (rc = launch(b, K_MOM_ASM, grid))) return rc;
        // folded: the pressure assembly forms u*, v* and finalises the finish
        if (!(b->fold & kFoldMom) && (rc = launch(b, K_MOM_FIN, grid))) return rc;
    }
    if ((rc = launch(b, K_PRES_ASM, grid))) return rc;
    if (b->anySOR)
        for (int k = 0; k < P; ++k)
            if ((rc = sorPair(b, grid))) return rc;
    if (!b->anyCG) {
        if ((rc = launch(b, K_CORR_OUTER, grid))) return rc;
        return dpMonitor(b);
    }
    if ((rc = mgSetup(b)) || (rc = mgVcycle(b))) return rc;
    for (int c = 0; c < R; ++c) {
        bool updateUnfinalised = false;
        for (int k = 0; k < P; ++k) {
            // with one refinement cycle x is formed late: every even
            // iteration from the third adds the two updates before it
            const bool foldX = (b->fold & kFoldPc) != 0 && k >= 2 && (k & 1) == 0;
            if ((rc = launchPcgQ(b, grid, k == 0, false, nullptr, foldX))) return rc;
            // the V-cycle after the last iteration would only feed a search
            // direction that is never used; before it, the update's stop
            // test is finalised together with the V-cycle's r.z
            if (k + 1 < P && b->mgLevels > 0) {
                const hipError_t e = launchWorkOnly(K_PCG_U, b->precision == CP_PREC_MIXED,
                                                    pRows(b->precision == CP_PREC_MIXED, b->tileRows), grid, b->stream,
                                                    kargsFor(b, K_PCG_U));
                if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
                if ((rc = mgVcycle(b, true))) return rc;
            } else if (k + 1 == P && b->precision == CP_PREC_MIXED) {
                // the last update's stop test is finalised with the
                // accumulation that follows it (PH_PCGU_REFACC)
                const hipError_t e = launchWorkOnly(K_PCG_U, true, pRows(true, b->tileRows), grid, b->stream,
                                                    kargsFor(b, K_PCG_U));
                if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
                updateUnfinalised = true;
            } else if ((rc = launch(b, K_PCG_U, grid))) {
                return rc;
            }
        }
        if (b->precision != CP_PREC_MIXED) break;
        if (b->fold & kFoldPc) {
            // p' = 0 + x is read by the correction from x itself (with the
            // CG updates k_pcg_q has not added to it yet): the
            // accumulation's bookkeeping only
            const hipError_t e =
                launchFinalize(updateUnfinalised ? PH_PCGU_REFACC : PH_REF_ACC, b->maxSlots, b->stream, kargs(b));
            if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
        } else if (!updateUnfinalised) {
            if ((rc = launch(b, K_REF_ACC, grid))) return rc;
        } else {
            KArgs ka = kargsFor(b, K_REF_ACC);
            hipError_t e = launchWorkOnly(K_REF_ACC, true, pRows(true, b->tileRows), grid, b->stream, ka);
            if (e == hipSuccess) e = launchFinalize(PH_PCGU_REFACC, b->maxSlots, b->stream, kargs(b));
            if (e != hipSuccess) return setError(&b->err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
        }
        if (c + 1 < R) {
            if ((rc = launch(b, K_REF_RES, grid))) return rc;
            if ((rc = mgVcycle(b))) return rc;
        }
    }
    if ((rc = launch(b, K_CORR_OUTER, grid))) return rc;
    return dpMonitor(b);
}

int graphRun(cp_batch* b, int steps) {
    std::string& err = b->err;
    const int J = maxOverSlots(b, &SlotDesc::maxMomIter);
    const int P = maxOverSlots(b, &SlotDesc::maxPresIter);
    const int R = maxOverSlots(b, &SlotDesc::refineCycles);
    if (!b->useGraph) {
        // the same fixed-budget sequence, launched kernel by kernel
        const int grid = gridFor(b, 1 << 30);
        for (int k = 0; k < steps; ++k) {
            const int rc = enqueueFixedStep(b, J, P, R, grid);
            if (rc) return rc;
        }
        return CP_OK;
    }
    const int families =
        b->anyExplicit | b->anyImplicit << 1 | b->anySOR << 2 | b->anyCG << 3 | b->anyMgCG << 4 | b->anyJacCG << 5;
    const int shapeL = ((b->mgLevels * (kMaxLevels + 1) + b->mgTail) * 2 + b->dpAny) * 64 + families;
    if (!b->graphExec || J != b->graphJ || P != b->graphP || R != b->graphR || shapeL != b->graphL) {
        if (b->graphExec) {
            // a replay of the old graph may still be in flight (a tiling
            // change that only sets the workgroups per CU does not
            // synchronise): let it finish before its executable goes
            HIP_TRY(hipStreamSynchronize(b->stream));
            HIP_TRY(hipGraphExecDestroy(b->graphExec));
            b->graphExec = nullptr;
        }
        const int grid = gridFor(b, 1 << 30);  // fixed: independent of the admitted slots
        hipGraph_t g;
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueueFixedStep(b, J, P, R, grid);
        hipGraph_t captured = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &captured);
        if (rc) return rc;
        if (e != hipSuccess) return setError(&err, CP_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
        g = captured;
        HIP_TRY(hipGraphInstantiate(&b->graphExec, g, nullptr, nullptr, 0));
        HIP_TRY(hipGraphDestroy(g));
        b->graphJ = J;
        b->graphP = P;
        b->graphR = R;
        b->graphL = shapeL;
    }
    for (int k = 0; k < steps; ++k) HIP_TRY(hipGraphLaunch(b->graphExec, b->stream));
    return CP_OK;
}

void fillStats(const SlotState& st, const SlotDesc& d, cp_step_stats* out) {
    if (!out) return;
    out->error = st.error;
    out->residU = st.residU;
    out->residV = st.residV;
    out->residMass = st.residMass;
    out->pseudoCFL = st.pseudoCFL;
    out->itU = st.itU;
    out->itV = st.itV;
    out->itP = st.itP;
    out->iter = st.iter;
    out->converged = (st.status == ST_CONVERGED) ? 1 : 0;
    out->health = st.health ? CP_E_DIVERGED : 0;
    (void)d;
}

int planeSample(cp_batch* b, int s, double x, double gmin, cp_plane_metrics* out) {
    std::string& err = b->err;
    if (!out) return setError(&err, CP_E_ARG, "null output");
    if (s < 0 || s >= b->maxSlots || !b->slots[s].used) return setError(&err, CP_E_ARG, "slot %d not in use", s);
    b->hPlaneReq[0].slot = s;
    b->hPlaneReq[0].x = x;
    b->hPlaneReq[0].gammaMin = gmin;
    HIP_TRY(hipMemcpyAsync(b->dPlaneReq, b->hPlaneReq, sizeof(PlaneReq), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(launchPlane(1, b->stream, b->dDesc, b->dPlaneReq, b->dPlaneOut));
    HIP_TRY(hipMemcpyAsync(b->hPlaneOut, b->dPlaneOut, sizeof(PlaneOut), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    const PlaneOut& o = b->hPlaneOut[0];
    out->flowArea = o.flowArea;
    out->massFlux = o.massFlux;
    out->avgStatic = o.avgStatic;
    out->avgDynamic = o.avgDynamic;
    out->avgTotal = o.avgTotal;
    out->valid = o.valid;
    return CP_OK;
}

int getFieldSlot(cp_batch* b, int s, int which, double* out) {
    std::string& err = b->err;
    if (!out) return setError(&err, CP_E_ARG, "null output");
    const SlotHost& sh = b->slots[s];
    const SlotDesc& d = b->hDesc[s];
    const int M = sh.M, N = sh.N, P = sh.pitch;
    switch (which) {
        case CP_F_U: return downloadPlane(b, d.u, P, M + 1, N, out);
        case CP_F_V: return downloadPlane(b, d.v, P, M, N + 1, out);
        case CP_F_P: return downloadPlane(b, d.p, P, M + 1, N + 1, out);
        case CP_F_USTAR: return downloadPlane(b, d.us, P, M + 1, N, out);
        case CP_F_VSTAR: return downloadPlane(b, d.vs, P, M, N + 1, out);
        case CP_F_PSTAR: return downloadPlane(b, d.pc, P, M + 1, N + 1, out);
        case CP_F_DE: return (d.dE32 ? downloadPlaneF : downloadPlane)(b, d.dE, P, M + 1, N, out);
        case CP_F_DN: return (d.dE32 ? downloadPlaneF : downloadPlane)(b, d.dN, P, M, N + 1, out);
        case CP_F_GAMMA:
        case CP_F_ALPHA: {
            // derived on the host from the parameters, as at admission
            std::vector<double> aU, aV, gam, g, al;
            std::vector<double> ct(size_t(M) * N, 0.0);
            // recover cellType from the stored gamma plane
            std::vector<double> gp(size_t(M + 1) * P);
            int rc = downloadPlane(b, d.gam, P, M + 1, P, gp.data());
            if (rc) return rc;
            // gam holds gamma after the wall-row forcing; alpha follows from it
            const double amax = alphaMaxOf(sh.prm);
            for (int i = 0; i < M; ++i)
                for (int j = 0; j < N; ++j) {
                    const double gv = gp[size_t(i + 1) * P + j + 1];
                    if (which == CP_F_GAMMA) out[size_t(i) * N + j] = gv;
                    else out[size_t(i) * N + j] = amax * (1.0 - gv) / (1.0 + sh.prm.brinkmanQ * gv + 1e-10);
                }
            if (which == CP_F_ALPHA && M >= 2)
                for (int j = 0; j < N; ++j) {
                    out[j] = amax;
                    out[size_t(M - 2) * N + j] = amax;
                }
            (void)ct;
            return CP_OK;
        }
        default: return setError(&err, CP_E_ARG, "unknown field %d", which);
    }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int cp_params_default(cp_params* p) {
    if (!p) return setError(nullptr, CP_E_ARG, "null params");
    cp_fill_reference_defaults(p);
    return CP_OK;
}

const char* cp_last_error(const void* ctx) {
    (void)ctx;
    std::lock_guard<std::mutex> lk(g_errMu);
    static thread_local std::string copy;
    copy = g_lastError;
    return copy.c_str();
}

int cp_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    int usable = 0;
    for (int d = 0; d < n; ++d) usable += isGfx950(d) ? 1 : 0;
    return usable;
}

// ---- batch -----------------------------------------------------------------
int cp_batch_create(int precision, size_t hbm_budget, int max_slots, cp_batch** out) {
    return batchCreate(precision, hbm_budget, max_slots, out, nullptr);
}

int cp_batch_destroy(cp_batch* b) {
    batchDestroy(b);
    return CP_OK;
}

size_t cp_batch_footprint(const cp_batch* b, int M, int N) { return b ? b->footprint(M, N) : 0; }

int cp_batch_admit(cp_batch* b, const cp_params* prm, const double* cellType, int* slot) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    return batchAdmit(b, prm, cellType, slot);
}

int cp_batch_run(cp_batch* b, int steps) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    if (steps < 0) return setError(&b->err, CP_E_ARG, "negative steps");
    if (b->totalTiles == 0 || steps == 0) return CP_OK;
    return graphRun(b, steps);
}

int cp_batch_poll(cp_batch* b, cp_slot_info* info) {
    if (!b || !info) return setError(nullptr, CP_E_ARG, "null argument");
    int rc = pollState(b);
    if (rc) return rc;
    for (int s = 0; s < b->maxSlots; ++s) {
        cp_slot_info& o = info[s];
        memset(&o, 0, sizeof(o));
        if (!b->slots[s].used) {
            o.status = CP_SLOT_FREE;
            continue;
        }
        const SlotState& st = b->hState[s];
        o.status = st.status == ST_RUNNING ? CP_SLOT_RUNNING
                   : st.status == ST_CONVERGED ? CP_SLOT_CONVERGED
                   : st.status == ST_MAXITER ? CP_SLOT_MAXITER
                                             : CP_SLOT_FAILED;
        o.iter = st.iter;
        o.error = st.error;
        o.residU = st.residU;
        o.residV = st.residV;
        o.residMass = st.residMass;
        o.health = st.health ? CP_E_DIVERGED : 0;
        o.M = b->slots[s].M;
        o.N = b->slots[s].N;
        o.itU = st.itU;
        o.itV = st.itV;
        o.itP = st.itP;
        o.pseudoCFL = st.pseudoCFL;
    }
    return CP_OK;
}

int cp_batch_get_fields(cp_batch* b, int slot, double* u, double* v, double* p) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    if (slot < 0 || slot >= b->maxSlots || !b->slots[slot].used)
        return setError(&b->err, CP_E_ARG, "slot %d not in use", slot);
    int rc;
    if (u && (rc = getFieldSlot(b, slot, CP_F_U, u))) return rc;
    if (v && (rc = getFieldSlot(b, slot, CP_F_V, v))) return rc;
    if (p && (rc = getFieldSlot(b, slot, CP_F_P, p))) return rc;
    return CP_OK;
}

int cp_batch_plane(cp_batch* b, int slot, double x, double gamma_min, cp_plane_metrics* out) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    return planeSample(b, slot, x, gamma_min, out);
}

int cp_batch_evict(cp_batch* b, int slot) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    return batchEvict(b, slot);
}

int cp_batch_set_tiling(cp_batch* b, int tile_rows, int wg_per_cu) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    if (!(tile_rows == 0 || tile_rows == 1 || tile_rows == 2 || tile_rows == 4 || tile_rows == 8) || wg_per_cu < 0 ||
        wg_per_cu > 64)
        return setError(&b->err, CP_E_ARG, "tiling out of range (rows %d, wg/CU %d)", tile_rows, wg_per_cu);
    if (tile_rows > 0 && tile_rows != b->tileRows) {
        // the tiling is only a decomposition of each slot's grid: retile all
        b->tileRows = tile_rows;
        for (int s = 0; s < b->maxSlots; ++s)
            if (b->slots[s].used) setTiling(b->hDesc[s], tile_rows, b->mgRowsEff());
        const int rc = rebuildTiles(b);
        if (rc) return rc;
    }
    if (wg_per_cu > 0) b->wgPerCU = wg_per_cu;
    b->graphJ = -1;
    return CP_OK;
}

constexpr int kProbeEvery = 10;     // polls between timing bursts
constexpr int kProbeLaunches = 16;  // rounds of launches per burst

namespace {
int burst(cp_batch* const* bs, int nb, Kern k, int rounds, double* ms, double* units, int* launches);
}

}  // extern "C"

namespace {

// The request queue of one cp_serve / cp_serve_pool call, shared by the
// batches that serve it (one host thread each, each batch on its own
// stream). Requests are taken first come, first served by whichever batch
// has room. The GPU gate lets a timing burst run with no other batch's work
// in flight: every worker holds it shared while it has device work queued
// (admission, step graph, poll, read-out), a burst holds it exclusively.
struct ServeQueue {
    cp_batch* const* batches = nullptr;  // every batch serving the queue
    int nbatches = 0;
    const cp_request* reqs = nullptr;
    cp_result* results = nullptr;
    int n = 0;
    std::vector<int> order;  // request indices by arrival
    std::mutex mu;           // admissions: next
    int next = 0;
    std::atomic<bool> abort{false};
    std::chrono::steady_clock::time_point t0;
    std::mutex gmu;
    std::condition_variable gcv;
    int active = 0;          // workers inside the shared gate
    bool exclusive = false;  // a burst waits for, or holds, the GPU alone

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    void enterShared() {
        std::unique_lock<std::mutex> lk(gmu);
        gcv.wait(lk, [&] { return !exclusive; });
        ++active;
    }
    void leaveShared() {
        std::lock_guard<std::mutex> lk(gmu);
        --active;
        gcv.notify_all();
    }
    void enterExclusive() {
        std::unique_lock<std::mutex> lk(gmu);
        gcv.wait(lk, [&] { return !exclusive; });
        exclusive = true;
        gcv.wait(lk, [&] { return active == 0; });
    }
    void leaveExclusive() {
        std::lock_guard<std::mutex> lk(gmu);
        exclusive = false;
        gcv.notify_all();
    }
};

// One batch's serve loop over the shared queue. Returns CP_OK when no
// request is left to admit and the batch has finished its own, or the first
// error (which stops the other workers at their next iteration).
int serveWorker(cp_batch* b, ServeQueue& q, int steps_per_poll) {
    if (hipSetDevice(b->device) != hipSuccess) return setError(&b->err, CP_E_HIP, "hipSetDevice failed");
    std::vector<int> reqOf(b->maxSlots, -1);
    b->probes.clear();
    long polls = 0;
    int running = 0;
    while (!q.abort.load()) {
        q.enterShared();
        int rc = CP_OK;
        {  // admissions into this batch, first come first served
            std::lock_guard<std::mutex> lk(q.mu);
            while (q.next < q.n && q.reqs[q.order[q.next]].arrival <= q.now()) {
                const int r = q.order[q.next];
                int slot = -1;
                const int arc = batchAdmit(b, q.reqs[r].prm, q.reqs[r].cellType, &slot);
                if (arc == CP_E_CAPACITY && running > 0) break;  // wait for an eviction here, or another batch
                ++q.next;
                if (arc) {  // invalid, or larger than the whole batch
                    q.results[r].rc = arc;
                    q.results[r].tAdmit = q.results[r].tDone = q.now();
                    continue;
                }
                q.results[r].slot = slot;
                q.results[r].tAdmit = q.now();
                reqOf[slot] = r;
                ++running;
            }
        }
        if (running == 0) {
            q.leaveShared();
            double wait = 0.0;
            {
                std::lock_guard<std::mutex> lk(q.mu);
                if (q.next >= q.n) return CP_OK;  // nothing left for this batch
                wait = q.reqs[q.order[q.next]].arrival - q.now();
            }
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
            continue;
        }
        rc = graphRun(b, steps_per_poll);
        if (!rc) rc = pollState(b);
        for (int s = 0; !rc && s < b->maxSlots; ++s) {
            const int r = reqOf[s];
            if (r < 0 || b->hState[s].status == ST_RUNNING) continue;
            cp_result& res = q.results[r];
            const SlotState& st = b->hState[s];
            res.status = st.status;
            res.iter = st.iter;
            res.error = st.error;
            if (st.health) res.rc = CP_E_DIVERGED;  // failed on a non-finite field; evicted like any other
            const double xOut = double(b->slots[s].N) * b->slots[s].prm.hx;
            cp_plane_metrics in, out;
            if ((rc = planeSample(b, s, 0.0, q.reqs[r].gammaMin, &in))) break;
            if ((rc = planeSample(b, s, xOut, q.reqs[r].gammaMin, &out))) break;
            res.dp = in.avgTotal - out.avgTotal;
            res.tDone = q.now();
            if ((rc = batchEvict(b, s))) break;
            reqOf[s] = -1;
            --running;
        }
        q.leaveShared();
        if (!rc && b->probeMask && polls++ % kProbeEvery == 0) {
            // timing bursts of the probed kernels, each launch round covering
            // the running slots of every batch of the queue in turn, with the
            // GPU to itself (the other workers' streams are idle)
            q.enterExclusive();
            for (int k = 0; !rc && k < 32; ++k) {
                if (!(b->probeMask & (1 << k))) continue;
                double ms = 0.0, units = 0.0;
                int launches = 0;
                rc = burst(q.batches, q.nbatches, Kern(k), kProbeLaunches, &ms, &units, &launches);
                if (!rc && launches > 0) b->probes.push_back({k, launches, ms, units});
            }
            q.leaveExclusive();
        }
        if (rc) {
            q.abort.store(true);
            return rc;
        }
    }
    return setError(&b->err, CP_E_STATE, "serve stopped: another batch of the pool failed");
}

int serveChecked(cp_batch* const* bs, int nb, const cp_request* reqs, int n, int steps_per_poll, cp_result* results) {
    for (int i = 0; i < nb; ++i) {
        std::string& err = bs[i]->err;
        if (n < 0 || (n > 0 && (!reqs || !results))) return setError(&err, CP_E_ARG, "null requests or results");
        if (steps_per_poll < 1) return setError(&err, CP_E_ARG, "steps_per_poll must be >= 1");
        for (int s = 0; s < bs[i]->maxSlots; ++s)
            if (bs[i]->slots[s].used) return setError(&err, CP_E_STATE, "cp_serve needs an empty batch");
    }
    ServeQueue q;
    q.batches = bs;
    q.nbatches = nb;
    q.reqs = reqs;
    q.results = results;
    q.n = n;
    q.order.resize(n);
    for (int k = 0; k < n; ++k) q.order[k] = k;
    std::stable_sort(q.order.begin(), q.order.end(), [&](int x, int y) { return reqs[x].arrival < reqs[y].arrival; });
    for (int k = 0; k < n; ++k) {
        results[k] = cp_result();
        results[k].slot = -1;
        results[k].status = CP_SLOT_FAILED;
    }
    q.t0 = std::chrono::steady_clock::now();
    std::vector<int> rcs(nb, CP_OK);
    std::vector<std::thread> th;
    for (int i = 1; i < nb; ++i) th.emplace_back([&, i] { rcs[i] = serveWorker(bs[i], q, steps_per_poll); });
    rcs[0] = serveWorker(bs[0], q, steps_per_poll);
    for (auto& t : th) t.join();
    // the first failure, not a worker stopped by it
    for (int i = 0; i < nb; ++i)
        if (rcs[i] != CP_OK && rcs[i] != CP_E_STATE) return rcs[i];
    for (int i = 0; i < nb; ++i)
        if (rcs[i] != CP_OK) return rcs[i];
    return CP_OK;
}

}  // namespace

extern "C" {

int cp_serve(cp_batch* b, const cp_request* reqs, int n, int steps_per_poll, cp_result* results) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    return serveChecked(&b, 1, reqs, n, steps_per_poll, results);
}

int cp_serve_pool(cp_batch* const* batches, int nbatches, const cp_request* reqs, int n, int steps_per_poll,
                  cp_result* results) {
    if (!batches || nbatches < 1) return setError(nullptr, CP_E_ARG, "no batches");
    for (int i = 0; i < nbatches; ++i) {
        if (!batches[i]) return setError(nullptr, CP_E_ARG, "null batch %d", i);
        for (int j = 0; j < i; ++j)
            if (batches[j] == batches[i]) return setError(nullptr, CP_E_ARG, "batch %d given twice", i);
        if (batches[i]->precision != batches[0]->precision)
            return setError(nullptr, CP_E_ARG, "batches of different precision");
    }
    return serveChecked(batches, nbatches, reqs, n, steps_per_poll, results);
}

int cp_batch_set_mg_tiling(cp_batch* b, int rows) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    const int maxRows = b->precision == CP_PREC_MIXED ? 8 : 2;
    if (!(rows == 0 || rows == 2 || rows == 4 || rows == 8) || rows > maxRows)
        return setError(&b->err, CP_E_ARG, "multigrid tile rows %d out of range (0, 2%s)", rows,
                        maxRows == 8 ? ", 4, 8" : "");
    b->mgTileRows = rows;
    for (int s = 0; s < b->maxSlots; ++s)
        if (b->slots[s].used) setTiling(b->hDesc[s], b->tileRows, b->mgRowsEff());
    const int rc = rebuildTiles(b);
    if (rc) return rc;
    b->graphJ = -1;
    return CP_OK;
}

int cp_batch_autotune(cp_batch* b, int steps, int* rows_out, int* mg_rows_out, int* wg_out, int* order_out,
                      double* ms_out) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    std::string& err = b->err;
    if (steps < 1 || steps > 1000) return setError(&err, CP_E_ARG, "steps must be in [1, 1000]");
    bool any = false;
    for (int s = 0; s < b->maxSlots; ++s) any = any || (b->slots[s].used && b->hState[s].status == ST_RUNNING);
    if (!any) return setError(&err, CP_E_STATE, "autotune needs running slots");
    const bool f32 = b->precision == CP_PREC_MIXED;
    const int rowsC[] = {1, 2, 4, 8};
    const int mgC[] = {2, 4, 8};
    const int wgC[] = {4, 8, 16};
    hipEvent_t ev[2] = {nullptr, nullptr};
    // one candidate: retile, capture, time `steps` replays
    auto timed = [&](int r, int m, int w, int order, double& per) -> int {
        b->tileRows = r;
        b->mgTileRows = m;
        b->wgPerCU = w;
        b->tileOrder = order;
        for (int s = 0; s < b->maxSlots; ++s)
            if (b->slots[s].used) setTiling(b->hDesc[s], r, b->mgRowsEff());
        int rc = rebuildTiles(b);
        if (rc) return rc;
        b->graphJ = -1;
        if ((rc = graphRun(b, 1))) return rc;  // captures the graph for this shape
        HIP_TRY(hipEventRecord(ev[0], b->stream));
        if ((rc = graphRun(b, steps))) return rc;
        HIP_TRY(hipEventRecord(ev[1], b->stream));
        HIP_TRY(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        per = double(ms) / steps;
        return CP_OK;
    };
    double best = 1e300;
    int bestR = b->tileRows, bestMg = b->mgTileRows, bestWg = b->wgPerCU, bestOrder = b->tileOrder;
    int rc = CP_OK;
    if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess)
        rc = setError(&err, CP_E_HIP, "hipEventCreate failed");
    // tile shapes in the batch's current order, then the orders with the
    // fastest shape
    const int order0 = b->tileOrder;
    for (int r : rowsC) {
        for (int m : mgC) {
            for (int w : wgC) {
                if (rc || m > (f32 ? 8 : 2)) continue;
                double per = 0.0;
                if ((rc = timed(r, m, w, order0, per))) break;
                if (per < best) {
                    best = per;
                    bestR = r;
                    bestMg = m;
                    bestWg = w;
                }
            }
        }
    }
    for (int order = 0; order < 8 && !rc; ++order) {
        if (order == order0) continue;
        double per = 0.0;
        if ((rc = timed(bestR, bestMg, bestWg, order, per))) break;
        if (per < best) {
            best = per;
            bestOrder = order;
        }
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    b->tileRows = bestR;
    b->mgTileRows = bestMg;
    b->wgPerCU = bestWg;
    b->tileOrder = bestOrder;
    for (int s = 0; s < b->maxSlots; ++s)
        if (b->slots[s].used) setTiling(b->hDesc[s], bestR, b->mgRowsEff());
    if ((rc = rebuildTiles(b))) return rc;
    b->graphJ = -1;
    if ((rc = pollState(b))) return rc;
    if (rows_out) *rows_out = bestR;
    if (mg_rows_out) *mg_rows_out = bestMg;
    if (wg_out) *wg_out = bestWg;
    if (order_out) *order_out = bestOrder;
    if (ms_out) *ms_out = best;
    return CP_OK;
}

int cp_batch_set_tile_order(cp_batch* b, int order) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    if (order < 0 || order > 7) return setError(&b->err, CP_E_ARG, "tile order %d out of range [0, 7]", order);
    b->tileOrder = order;
    b->graphJ = -1;  // the captured launches carry the order
    return CP_OK;
}

int cp_batch_set_graph(cp_batch* b, int use_graph) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    b->useGraph = use_graph ? 1 : 0;
    return CP_OK;
}

// Kernels the timing bursts can time: each recomputes its outputs from
// inputs no launch of it changes, between two outer iterations (k_pcg_q: the
// search direction and q = A p; k_mom_assemble: the momentum coefficients,
// r0 and the first sweep; the finest level's k_mg_up: the V-cycle result z;
// each of which the next outer iteration recomputes before any reader), so a
// burst leaves every solve as it was.
static bool probeable(int k) { return k == int(K_PCG_Q) || k == int(K_MOM_ASM) || k == int(K_MG_UP); }
static_assert(CP_K_MOM_ASM == K_MOM_ASM && CP_K_PCG_Q == K_PCG_Q && CP_K_MG_UP == K_MG_UP, "probe kernel ids");

int cp_batch_probe(cp_batch* b, int kernel_mask) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    if (kernel_mask & ~((1 << CP_K_PCG_Q) | (1 << CP_K_MOM_ASM) | (1 << CP_K_MG_UP)))
        return setError(&b->err, CP_E_ARG,
                        "timing bursts are defined for CP_K_PCG_Q, CP_K_MOM_ASM and CP_K_MG_UP only (mask 0x%x)",
                        kernel_mask);
    b->probeMask = kernel_mask;
    return CP_OK;
}

int cp_batch_probe_read(cp_batch* b, int* kernel, int* launches, double* ms, double* units, int cap) {
    if (!b || !ms) return setError(nullptr, CP_E_ARG, "null argument");
    int n = 0;
    for (; n < int(b->probes.size()) && n < cap; ++n) {
        const auto& r = b->probes[n];
        if (kernel) kernel[n] = r.kernel;
        if (launches) launches[n] = r.launches;
        ms[n] = r.ms;
        if (units) units[n] = r.units;
    }
    return n;
}

int cp_batch_probe_burst(cp_batch* b, int kernel, int launches, double* ms_per_launch) {
    if (!b || !ms_per_launch) return setError(nullptr, CP_E_ARG, "null argument");
    std::string& err = b->err;
    if (!probeable(kernel))
        return setError(&err, CP_E_ARG, "timing bursts are defined for CP_K_PCG_Q, CP_K_MOM_ASM and CP_K_MG_UP only");
    if (launches < 1 || launches > 4096) return setError(&err, CP_E_ARG, "launches must be in [1, 4096]");
    double ms = 0.0, units = 0.0;
    int n = 0;
    const int rc = burst(&b, 1, Kern(kernel), launches, &ms, &units, &n);
    if (rc) return rc;
    *ms_per_launch = n > 0 ? ms / n : 0.0;
    return CP_OK;
}

namespace {
// Work units of kernel k over a batch's running slots: p' unknowns of the CG
// slots for k_pcg_q, nodes of the (M+1) x (N+1) grid of the implicit-momentum
// slots for k_mom_assemble, p' unknowns of the multigrid slots whose finest
// level is a grid level (below the tail) for k_mg_up.
double burstUnits(const cp_batch* b, Kern k) {
    double u = 0.0;
    for (int s = 0; s < b->maxSlots; ++s) {
        if (!b->slots[s].used || b->hState[s].status != ST_RUNNING) continue;
        const SlotDesc& d = b->hDesc[s];
        if (k == K_PCG_Q && d.presType == 1) u += double(d.M - 1) * double(d.N - 2);
        if (k == K_MOM_ASM && d.momType == 1) u += double(d.M + 1) * double(d.N + 1);
        if (k == K_MG_UP && d.presType == 1 && d.precond == CP_PRECOND_AMG && d.mgTail > 0)
            u += double(d.M - 1) * double(d.N - 2);
    }
    return u;
}

// `rounds` rounds of back-to-back launches of kernel k, each round one
// launch per batch with running slots, all on the first batch's stream,
// between one pair of events: the time per launch is the kernel's duration
// plus the dispatch gap between consecutive launches on one queue, as a
// timing loop measures it. With several batches the rounds rotate over all
// their arrays, so the working set is the pool's, not one batch's. Every
// batch's stream must be idle.
int burst(cp_batch* const* bs, int nb, Kern k, int rounds, double* ms, double* units, int* launches) {
    cp_batch* b0 = bs[0];
    std::string& err = b0->err;
    while (b0->burstEv.size() < 2) {
        hipEvent_t ev = nullptr;
        HIP_TRY(hipEventCreate(&ev));
        b0->burstEv.push_back(ev);
    }
    double perRound = 0.0;
    int nl = 0;
    for (int i = 0; i < nb; ++i) {
        const double u = burstUnits(bs[i], k);
        perRound += u;
        nl += u > 0.0 ? 1 : 0;
    }
    *ms = 0.0;
    *units = 0.0;
    *launches = 0;
    if (nl == 0) return CP_OK;
    HIP_TRY(hipEventRecord(b0->burstEv[0], b0->stream));
    for (int r = 0; r < rounds; ++r) {
        for (int i = 0; i < nb; ++i) {
            cp_batch* b = bs[i];
            if (burstUnits(b, k) <= 0.0) continue;
            KArgs a = kargsFor(b, k);
            a.force = 1;
            const bool f32 = b->precision == CP_PREC_MIXED;
            int rows = pOnly(k) ? pRows(f32, b->tileRows) : b->tileRows;
            if (k == K_MG_UP) {  // the finest level's tile list, as launchMg sets it
                a.tileSlot = b->dMgTileSlot;
                a.nTiles = b->dMgNTiles;
                a.level = 0;
                a.order = (b->tileOrder >> 2) & 1;
                rows = b->mgRowsEff();
            }
            if (k == K_PCG_Q) {
                const int rc = launchPcgQ(b, gridFor(b, 1 << 30), false, true, b0->stream);
                if (rc) {
                    if (b != b0) err = b->err;
                    return rc;
                }
                continue;
            }
            const hipError_t e = launchWorkOnly(k, f32, rows, gridFor(b, 1 << 30), b0->stream, a);
            if (e != hipSuccess) return setError(&err, CP_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
        }
    }
    HIP_TRY(hipEventRecord(b0->burstEv[1], b0->stream));
    HIP_TRY(hipEventSynchronize(b0->burstEv[1]));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, b0->burstEv[0], b0->burstEv[1]));
    *ms = double(t);
    *units = perRound * rounds;
    *launches = nl * rounds;
    return CP_OK;
}
}  // namespace

// ---- single-geometry context ------------------------------------------------
int cp_create(const cp_params* prm, const double* cellType, cp_ctx** out) {
    std::string err;
    int rc = validate(prm, &err);
    if (rc) return rc;
    if (!out || !cellType) return setError(nullptr, CP_E_ARG, "null argument");
    cp_batch* b = nullptr;
    // arena sized for exactly this geometry
    const size_t budget = roundUp(slotMainBytes(prm->precision, prm->M, prm->N) +
                                      slotMgBytes(prm->precision, prm->M, prm->N) + 256 * 64,
                                  1 << 20);
    rc = batchCreate(prm->precision, budget, 1, &b, nullptr);
    if (rc) return rc;
    b->tileRows = defaultTileRows(prm->M, prm->N);
    int slot = -1;
    rc = batchAdmit(b, prm, cellType, &slot);
    if (rc) {
        batchDestroy(b);
        return rc;
    }
    cp_ctx* c = new cp_ctx();
    c->b = b;
    c->slot = slot;
    *out = c;
    return CP_OK;
}

int cp_destroy(cp_ctx* ctx) {
    if (!ctx) return CP_OK;
    batchDestroy(ctx->b);
    delete ctx;
    return CP_OK;
}

static int setFieldsSlot(cp_batch* b, int slot, const double* u, const double* v, const double* p) {
    if (slot < 0 || slot >= b->maxSlots || !b->slots[slot].used)
        return setError(&b->err, CP_E_ARG, "slot %d not in use", slot);
    const SlotHost& sh = b->slots[slot];
    const SlotDesc& d = b->hDesc[slot];
    const int M = sh.M, N = sh.N, P = sh.pitch;
    int rc;
    if (u) {
        if ((rc = uploadPlane(b, d.u, P, M + 1, N, u))) return rc;
        if ((rc = uploadPlane(b, d.us, P, M + 1, N, u))) return rc;
    }
    if (v) {
        if ((rc = uploadPlane(b, d.v, P, M, N + 1, v))) return rc;
        if ((rc = uploadPlane(b, d.vs, P, M, N + 1, v))) return rc;
    }
    if (p && (rc = uploadPlane(b, d.p, P, M + 1, N + 1, p))) return rc;
    return CP_OK;
}

int cp_set_fields(cp_ctx* ctx, const double* u, const double* v, const double* p) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    return setFieldsSlot(ctx->b, ctx->slot, u, v, p);
}

int cp_batch_set_fields(cp_batch* b, int slot, const double* u, const double* v, const double* p) {
    if (!b) return setError(nullptr, CP_E_ARG, "null batch");
    return setFieldsSlot(b, slot, u, v, p);
}

int cp_get_fields(cp_ctx* ctx, double* u, double* v, double* p) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    return cp_batch_get_fields(ctx->b, ctx->slot, u, v, p);
}

int cp_get_field(cp_ctx* ctx, int which, double* out) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    return getFieldSlot(ctx->b, ctx->slot, which, out);
}

static int ctxStats(cp_ctx* ctx, cp_step_stats* st) {
    if (!st) return CP_OK;
    int rc = pollState(ctx->b);
    if (rc) return rc;
    fillStats(ctx->b->hState[ctx->slot], ctx->b->hDesc[ctx->slot], st);
    return CP_OK;
}

static int ctxRunning(cp_ctx* ctx) {
    int rc = pollState(ctx->b);
    if (rc) return rc;
    if (ctx->b->hState[ctx->slot].status != ST_RUNNING)
        return setError(&ctx->b->err, CP_E_STATE, "solve is not running (status %d)", ctx->b->hState[ctx->slot].status);
    return CP_OK;
}

int cp_momentum(cp_ctx* ctx, cp_step_stats* st) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    int rc = ctxRunning(ctx);
    if (rc) return rc;
    if ((rc = polledMomentum(ctx->b))) return rc;
    return ctxStats(ctx, st);
}

int cp_pressure(cp_ctx* ctx, cp_step_stats* st) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    int rc = ctxRunning(ctx);
    if (rc) return rc;
    if ((rc = polledPressure(ctx->b))) return rc;
    if (!st) return CP_OK;
    if ((rc = pollState(ctx->b))) return rc;
    const SlotState& s = ctx->b->hState[ctx->slot];
    fillStats(s, ctx->b->hDesc[ctx->slot], st);
    st->itP = ctx->b->precision == CP_PREC_MIXED ? s.pitTotal : s.piters;
    return CP_OK;
}

int cp_correct(cp_ctx* ctx) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    int rc = ctxRunning(ctx);
    if (rc) return rc;
    return launch(ctx->b, K_CORR, gridFor(ctx->b, ctx->b->totalTiles));
}

int cp_step(cp_ctx* ctx, cp_step_stats* st) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    int rc = ctxRunning(ctx);
    if (rc) return rc;
    if ((rc = polledStep(ctx->b, true))) return rc;
    return ctxStats(ctx, st);
}

int cp_run(cp_ctx* ctx, int check_every, cp_step_stats* last) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    cp_batch* b = ctx->b;
    int rc = pollState(b);
    if (rc) return rc;
    (void)check_every;
    while (b->hState[ctx->slot].status == ST_RUNNING) {
        if ((rc = polledStep(b, true))) return rc;
        if ((rc = pollState(b))) return rc;
    }
    return ctxStats(ctx, last);
}

int cp_plane(cp_ctx* ctx, double x, double gamma_min, cp_plane_metrics* out) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    return planeSample(ctx->b, ctx->slot, x, gamma_min, out);
}

int cp_sync(cp_ctx* ctx) {
    if (!ctx) return setError(nullptr, CP_E_ARG, "null ctx");
    std::string& err = ctx->b->err;
    HIP_TRY(hipStreamSynchronize(ctx->b->stream));
    return CP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the solve that allocates nothing: cp_run_noalloc
// ---------------------------------------------------------------------------
// Device scratch of cp_run_noalloc: one single-slot batch (descriptor, slot
// state, tile lists, reduction partials, plane buffers and the slot's arena)
// is laid out in it per call. Every device holds its own copy for the life
// of the process. In host code `scratch` names no device memory: the host
// uses the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) char scratch[CP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchMu;
char* g_scratchPtr[kMaxScratchDevices] = {};

// Device address of `scratch` on device `dev` (current), looked up once.
int scratchOf(int dev, char** out) {
    if (dev < 0 || dev >= kMaxScratchDevices)
        return setError(nullptr, CP_E_NODEV, "device ordinal %d beyond the %d the scratch table covers", dev,
                        kMaxScratchDevices);
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return setError(nullptr, CP_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratchPtr[dev] = static_cast<char*>(ptr);
    }
    *out = g_scratchPtr[dev];
    return CP_OK;
}

// Byte offsets of a one-slot batch in the scratch (256-byte aligned parts).
struct ScratchLayout {
    size_t desc, state, nTiles, pnTiles, mgnTiles, tileSlot, ptileSlot, mgTileSlot, partials, planeReq, planeOut;
    size_t arena, arenaBytes, total, maxTiles;
};

ScratchLayout scratchLayout(const SlotDesc& d, size_t footprint) {
    ScratchLayout L;
    size_t mt = size_t(std::max(d.ntiles, d.pntiles));
    for (int l = 0; l < d.nLevels; ++l) mt = std::max(mt, size_t(d.lv[l].ntiles));
    L.maxTiles = mt;
    size_t off = 0;
    auto take = [&](size_t n) {
        const size_t o = off;
        off = roundUp(off + n, 256);
        return o;
    };
    L.desc = take(sizeof(SlotDesc));
    L.state = take(sizeof(SlotState));
    L.nTiles = take(sizeof(int));
    L.pnTiles = take(sizeof(int));
    L.mgnTiles = take(sizeof(int) * kMaxLevels);
    L.tileSlot = take(sizeof(int) * mt);
    L.ptileSlot = take(sizeof(int) * mt);
    L.mgTileSlot = take(sizeof(int) * mt * kMaxLevels);
    L.partials = take(sizeof(double) * kNumPartials * mt);
    L.planeReq = take(sizeof(PlaneReq) * 2);
    L.planeOut = take(sizeof(PlaneOut) * 2);
    L.arenaBytes = roundUp(footprint, 256);
    L.arena = take(L.arenaBytes);
    L.total = off;
    return L;
}

}  // namespace

extern "C" {

int cp_run_noalloc(const cp_params* prm, const double* cellType, double* u, double* v, double* p,
                   cp_step_stats* last) {
    std::string perr;
    int rc = validate(prm, &perr);
    if (rc) return rc;
    if (!cellType) return setError(nullptr, CP_E_ARG, "null cellType");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || !isGfx950(dev))
        return setError(nullptr, CP_E_NODEV, "the current HIP device is not a gfx950 (MI355X) device");
    // a single-slot batch exactly as cp_create makes it, with its device
    // parts in the scratch, its host mirrors in the object, and the null
    // stream
    cp_batch b;
    std::string& err = b.err;
    b.precision = prm->precision;
    b.maxSlots = 1;
    b.device = dev;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    b.numCU = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    b.stream = nullptr;
    b.tileRows = defaultTileRows(prm->M, prm->N);
    const SlotDesc probe = makeDesc(*prm, b.precision, b.tileRows, b.mgRowsEff());
    const ScratchLayout L = scratchLayout(probe, b.footprint(prm->M, prm->N));
    if (L.total > size_t(CP_NOALLOC_SCRATCH_BYTES))
        return setError(nullptr, CP_E_CAPACITY,
                        "cp_run_noalloc: a %d x %d grid needs %zu bytes of device scratch, more than the %zu "
                        "of CP_NOALLOC_SCRATCH_BYTES",
                        prm->M, prm->N, L.total, size_t(CP_NOALLOC_SCRATCH_BYTES));
    char* base = nullptr;
    if ((rc = scratchOf(dev, &base))) return rc;
    b.dDesc = reinterpret_cast<SlotDesc*>(base + L.desc);
    b.dState = reinterpret_cast<SlotState*>(base + L.state);
    b.dNTiles = reinterpret_cast<int*>(base + L.nTiles);
    b.dPNTiles = reinterpret_cast<int*>(base + L.pnTiles);
    b.dMgNTiles = reinterpret_cast<int*>(base + L.mgnTiles);
    b.dTileSlot = reinterpret_cast<int*>(base + L.tileSlot);
    b.dPTileSlot = reinterpret_cast<int*>(base + L.ptileSlot);
    b.dMgTileSlot = reinterpret_cast<int*>(base + L.mgTileSlot);
    b.dPartials = reinterpret_cast<double*>(base + L.partials);
    b.dPlaneReq = reinterpret_cast<PlaneReq*>(base + L.planeReq);
    b.dPlaneOut = reinterpret_cast<PlaneOut*>(base + L.planeOut);
    b.maxTiles = L.maxTiles;
    b.arenaBase = base + L.arena;
    b.arena.init(b.arenaBase, L.arenaBytes);
    b.hState = &b.staticState;
    b.hPlaneReq = b.staticPlaneReq;
    b.hPlaneOut = b.staticPlaneOut;
    memset(b.hState, 0, sizeof(SlotState));
    b.hDesc.assign(1, SlotDesc());
    memset(b.hDesc.data(), 0, sizeof(SlotDesc));
    b.slots.assign(1, SlotHost());
    // the previous call's contents stay in the scratch: the slot state and
    // the tile counts start from zero here; admission zeroes the arena
    // region and writes the descriptor and the tile lists
    const int zeros[kMaxLevels] = {};
    HIP_TRY(hipMemcpyAsync(b.dState, b.hState, sizeof(SlotState), hipMemcpyHostToDevice, b.stream));
    HIP_TRY(hipMemcpyAsync(b.dNTiles, zeros, sizeof(int), hipMemcpyHostToDevice, b.stream));
    HIP_TRY(hipMemcpyAsync(b.dPNTiles, zeros, sizeof(int), hipMemcpyHostToDevice, b.stream));
    HIP_TRY(hipMemcpyAsync(b.dMgNTiles, zeros, sizeof(zeros), hipMemcpyHostToDevice, b.stream));
    HIP_TRY(hipStreamSynchronize(b.stream));
    int slot = -1;
    if ((rc = batchAdmit(&b, prm, cellType, &slot))) return rc;
    if ((u || v || p) && (rc = setFieldsSlot(&b, slot, u, v, p))) return rc;
    // cp_run's loop
    if ((rc = pollState(&b))) return rc;
    while (b.hState[slot].status == ST_RUNNING) {
        if ((rc = polledStep(&b, true))) return rc;
        if ((rc = pollState(&b))) return rc;
    }
    if (u && (rc = getFieldSlot(&b, slot, CP_F_U, u))) return rc;
    if (v && (rc = getFieldSlot(&b, slot, CP_F_V, v))) return rc;
    if (p && (rc = getFieldSlot(&b, slot, CP_F_P, p))) return rc;
    fillStats(b.hState[slot], b.hDesc[slot], last);
    HIP_TRY(hipStreamSynchronize(b.stream));
    return CP_OK;
}

}  // extern "C"
