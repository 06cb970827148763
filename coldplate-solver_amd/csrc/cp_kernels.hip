// cp_kernels.hip — gfx950 kernels of the SIMPLE backend.
//
// Every work kernel is a persistent grid-stride loop over the batch's tile
// list (tile -> slot table), so one launch covers every running geometry of
// a batch, and the launch shape never depends on which slots are admitted
// (the sequence can be captured once in a hipGraph). Per-slot reductions are
// deterministic and fence-free: each tile writes its block-reduced partials,
// and k_finalize (one workgroup per slot, launched next) sums that slot's
// partials in tile order and runs the slot's scalar logic (stop tests, CG
// coefficients, CFL control). The kernel boundary makes the partials
// visible; no tile ever waits on another. The reduction order depends only
// on the slot's own shape, so a geometry gives the same result alone or
// packed with others.
//
// T is the solver type: double (FP64 mode) or float (mixed mode). State
// arrays (u, v, p, u*, v*, p', dE, dN, alpha faces, gamma, mass imbalance)
// are double in both modes; reductions always accumulate in double.
#include <hip/hip_runtime.h>

#include "cp_device.h"

namespace cpdev {

// Field arrays live in device global memory, but their pointers come from
// the slot descriptor as plain (generic) pointers; accessed as such, every
// load and store is a flat access, which also counts against the LDS wait
// counter and serialises with the LDS traffic of the halo tiles. G() states
// the address space, so the accesses compile to global loads and stores.
template <typename T>
using gp = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gp<T> G(T* p) {
    return (gp<T>)p;
}

// Element k of a field array, with the byte offset formed in 32 bits and
// added to the (workgroup-uniform) base: this compiles to a global access
// with a scalar base and a 32-bit vector offset, without 64-bit address
// arithmetic per lane. Every field array is far below 4 GiB.
template <typename T>
__device__ __forceinline__ T ld32(gp<const T> base, uint32_t k) {
    return *reinterpret_cast<gp<const T>>(reinterpret_cast<gp<const char>>(base) + k * uint32_t(sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st32(gp<T> base, uint32_t k, T v) {
    *reinterpret_cast<gp<T>>(reinterpret_cast<gp<char>>(base) + k * uint32_t(sizeof(T))) = v;
}
// C adjacent elements k .. k+C-1 (k a multiple of C: the vector is aligned)
// in one access, and the store of the vector's members whose mask is set
// (all at once when all are).
template <typename T, int C>
using Vec = T __attribute__((ext_vector_type(C)));
template <int C, typename T>
__device__ __forceinline__ Vec<T, C> ldv(gp<const T> base, uint32_t k) {
    return *reinterpret_cast<gp<const Vec<T, C>>>(reinterpret_cast<gp<const char>>(base) + k * uint32_t(sizeof(T)));
}
template <int C, typename T>
__device__ __forceinline__ void stv(gp<T> base, uint32_t k, Vec<T, C> v) {
    *reinterpret_cast<gp<Vec<T, C>>>(reinterpret_cast<gp<char>>(base) + k * uint32_t(sizeof(T))) = v;
}
// Non-temporal forms, for data a launch touches once and no later kernel
// reads before it has left the caches: the access bypasses the caches'
// normal allocation, so a stream larger than the Infinity Cache does not
// evict lines other waves still reuse (the stencils' halo rows), and stores
// allocate nothing. Over 12 x 1024x512 p' systems rotating through 526 MB
// the PCG stencil takes 31.0 us with non-temporal stores and coefficient
// loads against 40.0 us without (71 against 55 % of 8 TB/s).
template <int C, typename T>
__device__ __forceinline__ Vec<T, C> ldvNT(gp<const T> base, uint32_t k) {
    return __builtin_nontemporal_load(
        reinterpret_cast<gp<const Vec<T, C>>>(reinterpret_cast<gp<const char>>(base) + k * uint32_t(sizeof(T))));
}
template <int C, typename T>
__device__ __forceinline__ void stvNT(gp<T> base, uint32_t k, Vec<T, C> v) {
    __builtin_nontemporal_store(v, reinterpret_cast<gp<Vec<T, C>>>(reinterpret_cast<gp<char>>(base) + k * uint32_t(sizeof(T))));
}
template <typename T>
__device__ __forceinline__ void st32NT(gp<T> base, uint32_t k, T v) {
    __builtin_nontemporal_store(v, reinterpret_cast<gp<T>>(reinterpret_cast<gp<char>>(base) + k * uint32_t(sizeof(T))));
}
// Element store, non-temporal when NT (compile-time) or nt (uniform).
template <bool NT, typename T>
__device__ __forceinline__ void stH(gp<T> base, size_t k, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, &base[k]);
    else base[k] = v;
}
template <typename T>
__device__ __forceinline__ void stR(gp<T> base, size_t k, T v, bool nt) {
    if (nt) __builtin_nontemporal_store(v, &base[k]);
    else base[k] = v;
}
// Vector load, non-temporal when NT.
template <int C, bool NT, typename T>
__device__ __forceinline__ Vec<T, C> ldvH(gp<const T> base, uint32_t k) {
    if constexpr (NT) return ldvNT<C>(base, k);
    else return ldv<C>(base, k);
}

template <int C, bool NT = false, typename T>
__device__ __forceinline__ void stMaskedV(gp<T> base, uint32_t k, Vec<T, C> v, const bool (&m)[C]) {
    bool all = true;
#pragma unroll
    for (int c = 0; c < C; ++c) all &= m[c];
    if (all) {
        if constexpr (NT) stvNT<C>(base, k, v);
        else stv<C>(base, k, v);
    } else {
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (m[c]) {
                if constexpr (NT) st32NT(base, k + c, v[c]);
                else st32(base, k + c, v[c]);
            }
    }
}
// The same for two arrays at the same index under one branch: two separate
// masked stores keep both vectors live across both branches (28 more VGPRs
// in the PCG stencil at 4 x 4 values per lane).
// NT: non-temporal stores.
template <int C, bool NT = false, typename T>
__device__ __forceinline__ void stMaskedV2(gp<T> a, Vec<T, C> va, gp<T> b, Vec<T, C> vb, uint32_t k,
                                           const bool (&m)[C]) {
    bool all = true;
#pragma unroll
    for (int c = 0; c < C; ++c) all &= m[c];
    if (all) {
        if constexpr (NT) {
            stvNT<C>(a, k, va);
            stvNT<C>(b, k, vb);
        } else {
            stv<C>(a, k, va);
            stv<C>(b, k, vb);
        }
    } else {
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (m[c]) {
                if constexpr (NT) {
                    st32NT(a, k + c, va[c]);
                    st32NT(b, k + c, vb[c]);
                } else {
                    st32(a, k + c, va[c]);
                    st32(b, k + c, vb[c]);
                }
            }
    }
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// max that propagates NaN (a non-finite residual must reach the health check)
__device__ __forceinline__ double nanmax(double a, double b) { return (a != a || a > b) ? a : b; }

template <bool IsMax>
__device__ __forceinline__ double combine(double a, double b) {
    if constexpr (IsMax) return nanmax(a, b);
    else return a + b;
}

template <bool IsMax>
__device__ __forceinline__ double waveReduce(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = combine<IsMax>(v, __shfl_xor(v, off, 64));
    return v;
}

// Reduces NQ values over the 256-thread block; MaxMask bit q selects max for
// value q (else sum). Thread 0 writes the results to dst[0..NQ). `calls`
// counts the block's calls (every wave makes the same calls, in order):
// consecutive calls alternate between two LDS buffers, so a call's writes
// never meet thread 0's reads of the call before it — between two calls on
// one buffer lies the barrier of the call in between, which thread 0 reaches
// only after its reads.
template <int NQ, unsigned MaxMask>
__device__ __forceinline__ void blockReduceTo(double (&v)[NQ], double* dst, int& calls) {
    static_assert(NQ <= kNumPartials, "entries of one reduction");
    __shared__ double sh[2][kBlock / 64][kNumPartials];
    const int b = calls & 1;
    ++calls;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        v[q] = ((MaxMask >> q) & 1u) ? waveReduce<true>(v[q]) : waveReduce<false>(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) sh[b][w][q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double r = sh[b][0][q];
            for (int k = 1; k < kBlock / 64; ++k)
                r = ((MaxMask >> q) & 1u) ? nanmax(r, sh[b][k][q]) : r + sh[b][k][q];
            dst[q] = r;
        }
    }
}

// One reduction of a block that makes no other call of this form between
// barriers (the per-slot kernels).
template <int NQ, unsigned MaxMask>
__device__ __forceinline__ void blockReduceTo(double (&v)[NQ], double* dst) {
    int calls = 0;
    blockReduceTo<NQ, MaxMask>(v, dst, calls);
}

// Sum of one partial entry over a slot's tiles for a work kernel that needs
// the slot's reduction itself: thread t adds tiles t, t + 256, ... in order
// (one load per thread for up to 256 tiles), then the block sum. Every
// workgroup of a launch gets the same value bit for bit. The result is valid
// in every thread.
__device__ __forceinline__ double slotSumLean(const double* partials, int tile0, int ntiles) {
    __shared__ double res[1];
    double v = 0.0;
    for (int k = int(threadIdx.x); k < ntiles; k += kBlock) v += G(partials)[size_t(tile0 + k) * kNumPartials];
    double vv[1] = {v};
    __syncthreads();
    blockReduceTo<1, 0u>(vv, res);
    __syncthreads();
    return res[0];
}

// Sums (or maxes) the partials of one slot's tiles in tile order (ntiles >=
// 1); the result is valid in every thread of the block. Each thread issues its loads in
// groups of 8 before combining them, so the latency is paid once per group.
template <int NQ, unsigned MaxMask>
__device__ __forceinline__ void slotReduce(const double* partials, int tile0, int ntiles, double (&out)[NQ]) {
    __shared__ double res[kNumPartials];
    double v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = 0.0;
    constexpr int kG = 8;
    for (int base = 0; base < ntiles; base += kBlock * kG) {
        double x[kG][NQ];
#pragma unroll
        for (int g = 0; g < kG; ++g) {
            // clamped index: every load is issued, the ones past the end discarded
            const int k = base + g * kBlock + int(threadIdx.x);
            const size_t kc = size_t(tile0 + min(k, ntiles - 1)) * kNumPartials;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const double xv = G(partials)[kc + q];
                x[g][q] = (k < ntiles) ? xv : 0.0;
            }
        }
#pragma unroll
        for (int g = 0; g < kG; ++g)
#pragma unroll
            for (int q = 0; q < NQ; ++q) v[q] = ((MaxMask >> q) & 1u) ? nanmax(v[q], x[g][q]) : v[q] + x[g][q];
    }
    __syncthreads();
    blockReduceTo<NQ, MaxMask>(v, res);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = res[q];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// tile decoding
// ---------------------------------------------------------------------------
struct Tile {
    int s;   // slot
    int i;   // first row of this wave's R rows
    int j;   // column of this lane
    int i0;  // first row of the tile
    int j0;  // first column of the tile
};

// A tile is 64 columns x 4R rows (R = rows per wave, a compile-time launch
// parameter shared by every slot of a batch): wave w of the workgroup owns
// rows [i0 + w R, i0 + (w+1) R).
// Tile t of the p' list (a slot's ptile0 .. ptile0 + pntiles - 1): rows from
// 1, so the tiles hold no row outside the unknowns' range 1 .. M-1; 128
// columns, lane l holding columns j0 + 2l and j0 + 2l + 1 (tl.j is the first).
template <int R, int C>
__device__ __forceinline__ Tile decodePTile(const KArgs& a, int t) {
    Tile tl;
    tl.s = __builtin_amdgcn_readfirstlane(a.tileSlot[t]);
    const SlotDesc& d = a.desc[tl.s];
    const int lt = t - d.ptile0;
    const int ty = lt / d.ptilesX;
    const int tx = lt - ty * d.ptilesX;
    tl.i0 = 1 + ty * (kTileH * R);
    tl.j0 = tx * (kTileW * C);
    tl.i = tl.i0 + int(threadIdx.x >> 6) * R;
    tl.j = tl.j0 + C * int(threadIdx.x & 63);
    return tl;
}

template <int R>
__device__ __forceinline__ Tile decodeTile(const KArgs& a, int t) {
    Tile tl;
    // the slot is uniform over the workgroup: keep it (and every descriptor
    // field read through it) in scalar registers
    tl.s = __builtin_amdgcn_readfirstlane(a.tileSlot[t]);
    const SlotDesc& d = a.desc[tl.s];
    const int lt = t - d.tile0;
    const int ty = lt / d.tilesX;
    const int tx = lt - ty * d.tilesX;
    tl.i0 = ty * (kTileH * R);
    tl.j0 = tx * kTileW;
    tl.i = tl.i0 + int(threadIdx.x >> 6) * R;
    tl.j = tl.j0 + int(threadIdx.x & 63);
    return tl;
}

// Tile order of a launch (KArgs::order, uniform over the launch).
// 0, XCD-split: the hardware dispatches workgroup b to XCD b mod 8 (each XCD
// has its own L2), and launches use a multiple of 8 workgroups. Each XCD
// walks one contiguous range of the tile list, so tiles that share halo
// cache lines (east/west and north/south neighbours) meet in the same L2
// instead of fetching those lines once per XCD. This pays while the arrays
// stay in the Infinity Cache.
// 1, launch order: workgroup b takes tiles b, b + grid, ...; the tiles in
// flight at any moment cover one contiguous stretch of the arrays, which
// DRAM serves faster when the working set exceeds the Infinity Cache
// (halo lines fetched by a second XCD then come from that cache).
// Every tile does the same work and writes its own partials in either order,
// so results do not depend on it.
#define XCD_TILE_LOOP(t, nT)                                                                       \
    const bool xo_ = a.order == 0;                                                                 \
    const int xcd_ = xo_ ? int(blockIdx.x) % kNumXCD : 0;                                          \
    const int lb_ = xo_ ? int(blockIdx.x) / kNumXCD : int(blockIdx.x);                             \
    const int per_ = xo_ ? ((nT) + kNumXCD - 1) / kNumXCD : (nT);                                  \
    const int tEnd_ = min((nT), (xcd_ + 1) * per_);                                                \
    const int gx_ = xo_ ? int(gridDim.x) / kNumXCD : int(gridDim.x);                               \
    for (int t = xcd_ * per_ + lb_; t < tEnd_; t += gx_)

__device__ __forceinline__ double clampv(double x, double m) { return fmax(-m, fmin(m, x)); }

// Copies a slot's descriptor into LDS, every thread of the block taking a
// share of its 8-byte words, and synchronises the block. Per-slot kernels
// read the descriptor from there: read through the global pointer, each
// field would be reloaded (one dependent round trip) after every store the
// compiler cannot separate from it.
__device__ __forceinline__ const SlotDesc& descToLds(SlotDesc* dst, const SlotDesc* src) {
    constexpr int kWords = int(sizeof(SlotDesc) / sizeof(uint64_t));
    static_assert(sizeof(SlotDesc) % sizeof(uint64_t) == 0, "descriptor size");
    const auto gs = G(reinterpret_cast<const uint64_t*>(src));
    uint64_t* ld = reinterpret_cast<uint64_t*>(dst);
    for (int w = int(threadIdx.x); w < kWords; w += int(blockDim.x)) ld[w] = gs[w];
    __syncthreads();
    return *dst;
}

// Row-major index of (i, j) clamped into the (M+1) x P array: lanes outside
// the unknowns load from a valid address and their values are discarded, so
// every load of a tile is issued unconditionally, ahead of any use.
__device__ __forceinline__ size_t clampIdx(int i, int j, int M, int P) {
    return size_t(min(max(i, 0), M)) * P + min(max(j, 0), P - 1);
}
__device__ __forceinline__ uint32_t clampIdx32(int i, int j, int M, int P) {
    return uint32_t(min(max(i, 0), M)) * uint32_t(P) + uint32_t(min(max(j, 0), P - 1));
}
// the same for the first column j of a C-column vector (j and P multiples of
// C: the vector stays aligned)
template <int C>
__device__ __forceinline__ uint32_t clampVec32(int i, int j, int M, int P) {
    return uint32_t(min(max(i, 0), M)) * uint32_t(P) + uint32_t(min(max(j, 0), P - C));
}

// lane l receives lane l-1's value (west) / lane l+1's value (east); lanes 0
// and 63 receive their own, which only halo lanes consume. A DPP move with a
// whole-wave shift (wave_shr:1 = 0x138, wave_shl:1 = 0x130) does this in the
// vector ALU; a lane without a source keeps the old value, its own (bound_ctrl
// off). 8-byte values move as two dwords. Every caller runs with the whole
// wave active (the branches around the stencils are uniform per slot): a DPP
// move from an inactive lane would also leave the old value.
template <int Ctrl, typename T>
__device__ __forceinline__ T dppWaveShift(T x) {
    if constexpr (sizeof(T) == 4) {
        const int v = __builtin_bit_cast(int, x);
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(v, v, Ctrl, 0xf, 0xf, false));
    } else {
        static_assert(sizeof(T) == 8, "4- or 8-byte lanes");
        const unsigned long long v = __builtin_bit_cast(unsigned long long, x);
        const int lo = int(uint32_t(v)), hi = int(uint32_t(v >> 32));
        const uint32_t rl = uint32_t(__builtin_amdgcn_update_dpp(lo, lo, Ctrl, 0xf, 0xf, false));
        const uint32_t rh = uint32_t(__builtin_amdgcn_update_dpp(hi, hi, Ctrl, 0xf, 0xf, false));
        return __builtin_bit_cast(T, (static_cast<unsigned long long>(rh) << 32) | rl);
    }
}
template <typename T>
__device__ __forceinline__ T fromWest(T x) { return dppWaveShift<0x138>(x); }
template <typename T>
__device__ __forceinline__ T fromEast(T x) { return dppWaveShift<0x130>(x); }

// ---------------------------------------------------------------------------
// momentum assembly (momentum_solver.cpp:185-298, delta form)
// ---------------------------------------------------------------------------
// Second-order upwind deferred correction (convection.cpp:53-186). The index
// guards are the reference's, including its U/V asymmetry.
template <bool IsU, typename Ptr, typename F>
__device__ __forceinline__ F souCorrection(Ptr f, size_t k, int i, int j, int P, int M, int N, F Fe, F Fw, F Fn,
                                           F Fs) {
    F S = F(0);
    const bool eOK = IsU ? (j + 2 < N) : (j + 2 <= N);
    const bool wOK = IsU ? (j + 1 < N) : (j + 1 <= N);
    const bool nOK = IsU ? (i + 2 <= M) : (i + 2 < M);
    const bool sOK = IsU ? (i + 1 <= M) : (i + 1 < M);
    // the face differences in the precision of the field, the products in
    // that of the fluxes
    if (Fe >= F(0)) {
        if (j >= 2) S -= Fe * F((1.5 * f[k] - 0.5 * f[k - 1]) - f[k]);
    } else {
        if (eOK) S -= Fe * F((1.5 * f[k + 1] - 0.5 * f[k + 2]) - f[k + 1]);
    }
    if (Fw >= F(0)) {
        if (j >= 2) S += Fw * F((1.5 * f[k - 1] - 0.5 * f[k - 2]) - f[k - 1]);
    } else {
        if (wOK) S += Fw * F((1.5 * f[k] - 0.5 * f[k + 1]) - f[k]);
    }
    if (Fn >= F(0)) {
        if (i >= 1) S -= Fn * F((1.5 * f[k] - 0.5 * f[k - P]) - f[k]);
    } else {
        if (nOK) S -= Fn * F((1.5 * f[k + P] - 0.5 * f[k + 2 * size_t(P)]) - f[k + P]);
    }
    if (Fs >= F(0)) {
        if (i >= 2) S += Fs * F((1.5 * f[k - P] - 0.5 * f[k - 2 * size_t(P)]) - f[k - P]);
    } else {
        if (sOK) S += Fs * F((1.5 * f[k] - 0.5 * f[k + P]) - f[k]);
    }
    return S;
}

// Descriptor fields of the momentum rows, copied to registers once per tile
// (read through the descriptor, they would be reloaded after every store,
// since the compiler cannot rule out aliasing).
template <typename T>
struct MomConst {
    int M, N, pseudo, localPseudo, convScheme;
    double convScale, rho, hx, hy, De, Dn, timeStep, minSpeed, hChar, sinkDiag, vol, uvAlpha;
    gp<double> dE, dN;
    gp<float> dE32, dN32;  // the same arrays holding floats (mixed-precision implicit rows)
    gp<T> mu[M_COUNT], mv[M_COUNT];
    __device__ __forceinline__ explicit MomConst(const SlotDesc& d)
        : M(d.M), N(d.N), pseudo(d.pseudo), localPseudo(d.localPseudo), convScheme(d.convScheme),
          convScale(d.convScale), rho(d.rho), hx(d.hx), hy(d.hy), De(d.De), Dn(d.Dn), timeStep(d.timeStep),
          minSpeed(d.minSpeed), hChar(d.hChar), sinkDiag(d.sinkDiag), vol(d.vol), uvAlpha(d.uvAlpha), dE(G(d.dE)),
          dN(G(d.dN)), dE32(G(reinterpret_cast<float*>(d.dE))), dN32(G(reinterpret_cast<float*>(d.dN))) {
#pragma unroll
        for (int q = 0; q < M_COUNT; ++q) {
            mu[q] = G(static_cast<T*>(d.mu[q]));
            mv[q] = G(static_cast<T*>(d.mv[q]));
        }
    }
};

// Assembles one U (IsU) or V row at node k: writes the relaxed diagonal, the
// east and north face fluxes (the row's four neighbour coefficients follow
// from them and from the west and south rows' fluxes, MomArr), dE/dN and the
// initial residual r0 = b - A u of the Jacobi system (the sweeps iterate on
// the correction delta = x - u from delta = 0, which is the reference's
// x0 = u), and returns r0 as stored. With `sweep` the row also counts as the
// first Jacobi sweep: from delta = 0 the sweep's residual is r0 and its
// iterate delta_1 = 0 + r0 / aP (T arithmetic, the stored values), which is
// not stored: the next sweep and the momentum finish form it from r0 and aP.
// Without it (no sweeps) the first delta buffer, which the momentum finish
// then reads, is zeroed. u, v, p are read from the workgroup's LDS tile
// (index ks, row pitch P); k indexes the node's global arrays.
template <typename T, bool IsU>
__device__ __forceinline__ T momRow(const MomConst<T>& d, const double* u, const double* v, const double* pp,
                                    size_t ks, int P, size_t k, int i, int j, double maxVel, double cfl,
                                    double alphaLocal, bool sweep) {
    double ue, uw, vn, vs;
    if (IsU) {
        ue = clampv(0.5 * (u[ks] + u[ks + 1]), maxVel);
        uw = clampv(0.5 * (u[ks - 1] + u[ks]), maxVel);
        vn = clampv(0.5 * (v[ks] + v[ks + 1]), maxVel);
        vs = clampv(0.5 * (v[ks - P] + v[ks - P + 1]), maxVel);
    } else {
        ue = clampv(0.5 * (u[ks] + u[ks + P]), maxVel);
        uw = clampv(0.5 * (u[ks - 1] + u[ks + P - 1]), maxVel);
        vn = clampv(0.5 * (v[ks] + v[ks + P]), maxVel);
        vs = clampv(0.5 * (v[ks - P] + v[ks]), maxVel);
    }
    const double cs = d.convScale;
    const double Fe = cs * d.rho * d.hy * ue, Fw = cs * d.rho * d.hy * uw;
    const double Fn = cs * d.rho * d.hx * vn, Fs = cs * d.rho * d.hx * vs;
    const double aE = d.De + fmax(0.0, -Fe);
    const double aW = d.De + fmax(0.0, Fw);
    const double aN = d.Dn + fmax(0.0, -Fn);
    const double aS = d.Dn + fmax(0.0, Fs);
    const double sumA = aE + aW + aN + aS;
    const auto f = IsU ? u : v;
    const double cur = f[ks];
    double trans = 0.0;
    if (d.pseudo) {
        double dt = fmax(d.timeStep, 1e-20);
        if (d.localPseudo) {
            const double speed = fmax(fabs(cur), d.minSpeed);
            dt = fmin(dt, cfl * d.hChar / speed);
            dt = fmax(dt, 1e-20);
        }
        trans = d.rho * d.vol / dt;
    }
    double aP0 = sumA + (Fe - Fw + Fn - Fs) + trans + d.sinkDiag + alphaLocal * d.vol;
    aP0 = fmax(aP0, 1e-10);
    const double aP = aP0 / d.uvAlpha;
    double Sp = IsU ? (pp[ks] - pp[ks + 1]) * d.hy : (pp[ks] - pp[ks + P]) * d.hx;
    Sp += trans * cur;  // uOld == u at momentum time (iterations.cpp:118-119)
    Sp += (1.0 - d.uvAlpha) / d.uvAlpha * aP0 * cur;
    if (d.convScheme == 1) Sp += souCorrection<IsU>(f, ks, i, j, P, d.M, d.N, Fe, Fw, Fn, Fs);
    if (IsU) d.dE[k] = d.hy / aP0;
    else d.dN[k] = d.hx / aP0;
    const double r0 = Sp + aE * f[ks + 1] + aW * f[ks - 1] + aN * f[ks + P] + aS * f[ks - P] - aP * cur;
    const gp<T>* A = IsU ? d.mu : d.mv;
    A[M_AP][k] = T(aP);
    A[M_FE][k] = T(Fe);
    A[M_FN][k] = T(Fn);
    const T r0t = T(r0);
    A[M_R0][k] = r0t;
    if (!sweep) A[M_D0][k] = T(0);
    return r0t;
}

// The mixed-precision form of momRow (T = float). Everything that enters
// the residual r0 = b - A u is computed in double, as in the FP64 row, so
// that delta = 0 (the outer fixed point) means the momentum equations hold
// in double precision: coefficients rounded to single precision there would
// move with every last-bit change of u and leave a noise floor of ~1e-8 in
// the converged velocities. What changes is how: the per-node divisions of
// the FP64 row become per-slot constants, 1/dt = max(1/timeStep, speed /
// (cfl hChar)) for the reference's dt = min(timeStep, cfl hChar / speed) and
// aP = aP0 / uvAlpha by the reciprocal of uvAlpha, and dE = hy / aP0, which
// only steers the correction and is 0 in effect at the fixed point, takes a
// single-precision reciprocal and is stored as a float. The stored diagonal
// and face fluxes are single precision.
struct MomScalars {
    double invTimeStep, invCflH, invUvAlpha, relaxK, rhoVol;
    float hy, hx;
};

template <bool IsU>
__device__ __forceinline__ float momRowMixed(const MomConst<float>& d, const MomScalars& m, const double* u,
                                             const double* v, const double* pp, size_t ks, int P, size_t k, int i,
                                             int j, double maxVel, double alphaLocal, bool sweep) {
    double ue, uw, vn, vs;
    if (IsU) {
        ue = clampv(0.5 * (u[ks] + u[ks + 1]), maxVel);
        uw = clampv(0.5 * (u[ks - 1] + u[ks]), maxVel);
        vn = clampv(0.5 * (v[ks] + v[ks + 1]), maxVel);
        vs = clampv(0.5 * (v[ks - P] + v[ks - P + 1]), maxVel);
    } else {
        ue = clampv(0.5 * (u[ks] + u[ks + P]), maxVel);
        uw = clampv(0.5 * (u[ks - 1] + u[ks + P - 1]), maxVel);
        vn = clampv(0.5 * (v[ks] + v[ks + P]), maxVel);
        vs = clampv(0.5 * (v[ks - P] + v[ks]), maxVel);
    }
    const double cs = d.convScale;
    const double Fe = cs * d.rho * d.hy * ue, Fw = cs * d.rho * d.hy * uw;
    const double Fn = cs * d.rho * d.hx * vn, Fs = cs * d.rho * d.hx * vs;
    const double aE = d.De + fmax(0.0, -Fe);
    const double aW = d.De + fmax(0.0, Fw);
    const double aN = d.Dn + fmax(0.0, -Fn);
    const double aS = d.Dn + fmax(0.0, Fs);
    const auto f = IsU ? u : v;
    const double cur = f[ks];
    double trans = 0.0;
    if (d.pseudo) {
        double invDt = m.invTimeStep;
        if (d.localPseudo) invDt = fmax(invDt, fmax(fabs(cur), d.minSpeed) * m.invCflH);
        trans = m.rhoVol * fmin(invDt, 1e20);
    }
    double aP0 = (aE + aW + aN + aS) + (Fe - Fw + Fn - Fs) + trans + d.sinkDiag + alphaLocal * d.vol;
    aP0 = fmax(aP0, 1e-10);
    const double aP = aP0 * m.invUvAlpha;
    double Sp = IsU ? (pp[ks] - pp[ks + 1]) * d.hy : (pp[ks] - pp[ks + P]) * d.hx;
    Sp += (trans + m.relaxK * aP0) * cur;  // uOld == u at momentum time (iterations.cpp:118-119)
    if (d.convScheme == 1) Sp += souCorrection<IsU>(f, ks, i, j, P, d.M, d.N, Fe, Fw, Fn, Fs);
    const float rcpAP0 = __builtin_amdgcn_rcpf(float(aP0));
    // the row's outputs are stored non-temporally (each is written once per
    // outer iteration; over the served pool the assembly takes 1.2 % of a
    // step less, and one batch alone 0.7 %)
    if (IsU) stH<true>(d.dE32, k, m.hy * rcpAP0);  // a float value: stored as one (SlotDesc::dE32)
    else stH<true>(d.dN32, k, m.hx * rcpAP0);
    const double r0 = Sp + aE * f[ks + 1] + aW * f[ks - 1] + aN * f[ks + P] + aS * f[ks - P] - aP * cur;
    const gp<float>* A = IsU ? d.mu : d.mv;
    stH<true>(A[M_AP], k, float(aP));
    stH<true>(A[M_FE], k, float(Fe));
    stH<true>(A[M_FN], k, float(Fn));
    const float r0t = float(r0);
    stH<true>(A[M_R0], k, r0t);
    if (!sweep) A[M_D0][k] = 0.0f;
    return r0t;
}

// Occupancy hint of the momentum row kernels: without it the compiler keeps
// 197 VGPRs (two waves per SIMD); at three waves it fits 168 without spills,
// and the assembly of 8 x 1024x512 runs in 118 instead of 134 us. Eight-row
// tiles need more registers than three waves allow and keep one.
constexpr int kMomWavesPerSimd = 3;
template <int R>
constexpr int momWaves() { return R >= 8 ? 1 : kMomWavesPerSimd; }

// The workgroup stages u, v and p of its tile with a two-deep halo (the SOU
// stencil reaches two nodes out) in LDS, so every neighbour read of the
// momentum rows is an LDS read; nodes outside the grid hold 0, and no row
// reads them. The kernel also makes the first Jacobi sweep of both systems
// (momRow) and leaves its per-tile residual partials, summed per lane in row
// order and block-reduced as k_jacobi does, so the finalisation after it is
// the first sweep's stop test.
template <typename T, int R>
__global__ __launch_bounds__(kBlock, momWaves<R>()) void k_mom_assemble(KArgs a) {
    constexpr int TH = kTileH * R, HH = TH + 4, HW = kTileW + 4;
    __shared__ double su[HH * HW], sv[HH * HW], sp[HH * HW];
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodeTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        if ((st.status != ST_RUNNING) | (d.momType != 1)) continue;
        const int M = d.M, N = d.N, P = d.pitch, j = tl.j;
        const bool sweep = d.maxMomIter > 0;
        const double maxVel = 3.0 * fmax(fabs(st.inletVelocity), 0.1);
        const double cfl = st.pseudoCFL;
        const auto gu = G(d.u), gv = G(d.v), gpp = G(d.p);
        const MomConst<T> c(d);
        MomScalars ms;
        if constexpr (sizeof(T) == 4) {
            ms.invTimeStep = 1.0 / fmax(d.timeStep, 1e-20);
            ms.invCflH = 1.0 / (cfl * d.hChar);
            ms.invUvAlpha = 1.0 / d.uvAlpha;
            ms.relaxK = (1.0 - d.uvAlpha) / d.uvAlpha;
            ms.rhoVol = d.rho * d.vol;
            ms.hy = float(d.hy);
            ms.hx = float(d.hx);
        }
        // every staging load of the tile first (clamped addresses), then the
        // LDS stores with the out-of-grid nodes zeroed
        constexpr int NE = HH * HW, NIT = (NE + kBlock - 1) / kBlock;
        double lu[NIT], lv[NIT], lp[NIT];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int idx = int(threadIdx.x) + it * kBlock;
            const int h = idx / HW, w = idx - h * HW;
            const size_t k = clampIdx(tl.i0 - 2 + h, tl.j0 - 2 + w, M, P);
            lu[it] = gu[k];
            lv[it] = gv[k];
            lp[it] = gpp[k];
        }
        // the Brinkman faces: floats in mixed-precision slots (SlotDesc::dE32,
        // set for every implicit slot of a single-precision batch)
        double aUr[R], aVr[R];
        if constexpr (sizeof(T) == 4) {
            const auto fU = G(reinterpret_cast<const float*>(d.aU)), fV = G(reinterpret_cast<const float*>(d.aV));
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const size_t k = clampIdx(tl.i + r, j, M, P);
                aUr[r] = fU[k];
                aVr[r] = fV[k];
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const size_t k = clampIdx(tl.i + r, j, M, P);
                aUr[r] = G(d.aU)[k];
                aVr[r] = G(d.aV)[k];
            }
        }
        __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int idx = int(threadIdx.x) + it * kBlock;
            if (NE % kBlock != 0 && idx >= NE) break;
            const int h = idx / HW, w = idx - h * HW;
            const int i = tl.i0 - 2 + h, jj = tl.j0 - 2 + w;
            const bool in = i >= 0 && i <= M && jj >= 0 && jj <= N;
            su[idx] = in ? lu[it] : 0.0;
            sv[idx] = in ? lv[it] : 0.0;
            sp[idx] = in ? lp[it] : 0.0;
        }
        __syncthreads();
        const int lane = int(threadIdx.x & 63), wv = int(threadIdx.x >> 6);
        double v[2] = {0.0, 0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            if (i > M || j > N) continue;
            const size_t k = size_t(i) * P + j;
            const size_t ks = size_t(wv * R + r + 2) * HW + (lane + 2);
            if (i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2) {
                T res;
                if constexpr (sizeof(T) == 4) res = momRowMixed<true>(c, ms, su, sv, sp, ks, HW, k, i, j, maxVel, aUr[r], sweep);
                else res = momRow<T, true>(c, su, sv, sp, ks, HW, k, i, j, maxVel, cfl, aUr[r], sweep);
                v[0] += double(res) * double(res);
            }
            if (i >= 1 && i <= M - 2 && j >= 1 && j <= N - 1) {
                T res;
                if constexpr (sizeof(T) == 4) res = momRowMixed<false>(c, ms, su, sv, sp, ks, HW, k, i, j, maxVel, aVr[r], sweep);
                else res = momRow<T, false>(c, su, sv, sp, ks, HW, k, i, j, maxVel, cfl, aVr[r], sweep);
                v[1] += double(res) * double(res);
            }
        }
        blockReduceTo<2, 0u>(v, a.partials + size_t(t) * kNumPartials, nred);
        // the workgroup holding the slot's first tile resets the Jacobi state
        // (no tile of this kernel reads it)
        if (t == d.tile0 && threadIdx.x == 0) {
            const int cap = d.maxMomIter;
            st.jk[0] = st.jk[1] = 0;
            st.jiters[0] = st.jiters[1] = 0;
            st.jdone[0] = st.jdone[1] = (cap <= 0) ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------------------
// Explicit momentum (iterations.cpp:176-441, momentumSolverType 0): every U
// and V unknown updated once from the previous iterate, with the implicit
// path's coefficients, no linear solve.
// ---------------------------------------------------------------------------
// One explicit point update at node k (LDS index ks, row pitch P): writes
// dE or dN and returns the clamped new velocity.
template <typename T, bool IsU>
__device__ __forceinline__ double momRowExplicit(const MomConst<T>& d, const double* u, const double* v,
                                                 const double* pp, size_t ks, int P, size_t k, int i, int j,
                                                 double maxVel, double cfl, double alphaLocal) {
    double ue, uw, vn, vs;
    if (IsU) {
        ue = clampv(0.5 * (u[ks] + u[ks + 1]), maxVel);
        uw = clampv(0.5 * (u[ks - 1] + u[ks]), maxVel);
        vn = clampv(0.5 * (v[ks] + v[ks + 1]), maxVel);
        vs = clampv(0.5 * (v[ks - P] + v[ks - P + 1]), maxVel);
    } else {
        ue = clampv(0.5 * (u[ks] + u[ks + P]), maxVel);
        uw = clampv(0.5 * (u[ks - 1] + u[ks + P - 1]), maxVel);
        vn = clampv(0.5 * (v[ks] + v[ks + P]), maxVel);
        vs = clampv(0.5 * (v[ks - P] + v[ks]), maxVel);
    }
    const double cs = d.convScale;
    const double Fe = cs * d.rho * d.hy * ue, Fw = cs * d.rho * d.hy * uw;
    const double Fn = cs * d.rho * d.hx * vn, Fs = cs * d.rho * d.hx * vs;
    const double aE = d.De + fmax(0.0, -Fe);
    const double aW = d.De + fmax(0.0, Fw);
    const double aN = d.Dn + fmax(0.0, -Fn);
    const double aS = d.Dn + fmax(0.0, Fs);
    const auto f = IsU ? u : v;
    const double cur = f[ks];
    const double Sdc = d.convScheme == 1 ? souCorrection<IsU>(f, ks, i, j, P, d.M, d.N, Fe, Fw, Fn, Fs) : 0.0;
    const double sumA = aE + aW + aN + aS;
    // computeLocalDt (iterations.cpp:210-216): no pseudo-time is an infinite step
    double trans = 0.0;
    if (d.pseudo) {
        double dt = d.timeStep;
        if (d.localPseudo) dt = fmin(d.timeStep, cfl * d.hChar / fmax(fabs(cur), d.minSpeed));
        trans = d.rho * d.vol / dt;
    }
    double aP0 = sumA + (Fe - Fw + Fn - Fs) + trans + d.sinkDiag + alphaLocal * d.vol;
    aP0 = fmax(aP0, 1e-10);
    const double dDenom = fmax(aP0, 1e-12);
    if (IsU) d.dE[k] = d.hy / dDenom;
    else d.dN[k] = d.hx / dDenom;
    double Sp = IsU ? (pp[ks] - pp[ks + 1]) * d.hy : (pp[ks] - pp[ks + P]) * d.hx;
    Sp += trans * cur;  // uOld == u here
    Sp += Sdc;
    const double aP = aP0 / d.uvAlpha;
    Sp += (1.0 - d.uvAlpha) / d.uvAlpha * aP0 * cur;
    const double nv = (aE * f[ks + 1] + aW * f[ks - 1] + aN * f[ks + P] + aS * f[ks - P] + Sp) / aP;
    return clampv(nv, maxVel);
}

// u* and v* of the unknowns, dE, dN, and per tile max |new - old| of U and V
// (NaN ignored, as the reference's comparisons do) and a NaN flag. The
// workgroup holding the slot's first tile marks the Jacobi state done.
template <typename T, int R>
__global__ __launch_bounds__(kBlock, momWaves<R>()) void k_mom_explicit(KArgs a) {
    constexpr int TH = kTileH * R, HH = TH + 4, HW = kTileW + 4;
    __shared__ double su[HH * HW], sv[HH * HW], sp[HH * HW];
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodeTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        if ((st.status != ST_RUNNING) | (d.momType != 0)) continue;
        const int M = d.M, N = d.N, P = d.pitch, j = tl.j;
        const double maxVel = 3.0 * fmax(fabs(st.inletVelocity), 0.1);
        const double cfl = st.pseudoCFL;
        const auto gu = G(d.u), gv = G(d.v), gpp = G(d.p);
        const auto us = G(d.us), vs = G(d.vs);
        const MomConst<T> c(d);
        constexpr int NE = HH * HW, NIT = (NE + kBlock - 1) / kBlock;
        double lu[NIT], lv[NIT], lp[NIT];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int idx = int(threadIdx.x) + it * kBlock;
            const int h = idx / HW, w = idx - h * HW;
            const size_t k = clampIdx(tl.i0 - 2 + h, tl.j0 - 2 + w, M, P);
            lu[it] = gu[k];
            lv[it] = gv[k];
            lp[it] = gpp[k];
        }
        double aUr[R], aVr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const size_t k = clampIdx(tl.i + r, j, M, P);
            aUr[r] = G(d.aU)[k];
            aVr[r] = G(d.aV)[k];
        }
        __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int idx = int(threadIdx.x) + it * kBlock;
            if (NE % kBlock != 0 && idx >= NE) break;
            const int h = idx / HW, w = idx - h * HW;
            const int i = tl.i0 - 2 + h, jj = tl.j0 - 2 + w;
            const bool in = i >= 0 && i <= M && jj >= 0 && jj <= N;
            su[idx] = in ? lu[it] : 0.0;
            sv[idx] = in ? lv[it] : 0.0;
            sp[idx] = in ? lp[it] : 0.0;
        }
        __syncthreads();
        const int lane = int(threadIdx.x & 63), wv = int(threadIdx.x >> 6);
        double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            if (i > M || j > N) continue;
            const size_t k = size_t(i) * P + j;
            const size_t ks = size_t(wv * R + r + 2) * HW + (lane + 2);
            if (i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2) {
                const double nv = momRowExplicit<T, true>(c, su, sv, sp, ks, HW, k, i, j, maxVel, cfl, aUr[r]);
                const double df = fabs(nv - su[ks]);
                if (df > v[0]) v[0] = df;
                if (!(nv == nv)) v[2] = 1.0;
                us[k] = nv;
            }
            if (i >= 1 && i <= M - 2 && j >= 1 && j <= N - 1) {
                const double nv = momRowExplicit<T, false>(c, su, sv, sp, ks, HW, k, i, j, maxVel, cfl, aVr[r]);
                const double df = fabs(nv - sv[ks]);
                if (df > v[1]) v[1] = df;
                if (!(nv == nv)) v[2] = 1.0;
                vs[k] = nv;
            }
        }
        blockReduceTo<3, 7u>(v, a.partials + size_t(t) * kNumPartials, nred);
        if (t == d.tile0 && threadIdx.x == 0) {  // no tile of this launch reads it
            st.jk[0] = st.jk[1] = 0;
            st.jiters[0] = st.jiters[1] = 0;
            st.jdone[0] = st.jdone[1] = 1;
        }
    }
}

// Velocity BCs on (u*, v*) of explicit slots (boundaries.cpp:48-106): walls,
// inlet, and the zero-gradient outlet from the unknowns' new values.
template <int R>
__global__ __launch_bounds__(kBlock) void k_mom_expl_bc(KArgs a) {
    const int nT = *a.nTiles;
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodeTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        const SlotState& st = a.state[tl.s];
        if ((st.status != ST_RUNNING) | (d.momType != 0)) continue;
        const int M = d.M, N = d.N, P = d.pitch, j = tl.j;
        const double Uin = st.inletVelocity;
        const auto us = G(d.us), vs = G(d.vs), gam = G(d.gam);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            if (i > M || j > N) continue;
            const size_t k = size_t(i) * P + j;
            if (j <= N - 1) {  // u* : (M+1) x N
                if (i == 0 || i == M) us[k] = 0.0;
                else if (j == 0) us[k] = (gam[k + 1] > 0.5) ? Uin : 0.0;  // gamma(i-1, 0) > 0.5
                else if (j == N - 1) us[k] = us[k - 1];
            }
            if (i <= M - 1) {  // v* : M x (N+1)
                if (i == 0 || i == M - 1 || j == 0) vs[k] = 0.0;
                else if (j == N) vs[k] = vs[k - 1];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Jacobi sweep (jacobi_solver.cpp:44-67): r = r0 - A delta,
// delta' = delta + r / aP, stop after the sweep where sum r^2 < tol^2 n.
// U and V are independent systems swept in the same launch, each with its
// own stop test and sweep count.
// ---------------------------------------------------------------------------
// Value of array `f` at (i, j) of a (M+1) x pitch plane, 0 outside the
// node range [0, M] x [0, jmax].
template <typename T>
__device__ __forceinline__ T at(const T* f, int i, int j, int M, int jmax, int P) {
    return (i >= 0 && i <= M && j >= 0 && j <= jmax) ? f[size_t(i) * P + j] : T(0);
}

// Jacobi correction of one system at node (i, j) for the virtual first
// iterate delta_1 = r0 / aP (the assembly's sweep, not stored): 0 outside the
// system's unknowns i in [1, iMax], j in [1, jMax], whose r0 and aP the
// caller loaded from clamped addresses.
template <typename T>
__device__ __forceinline__ T delta1(T r0, T aP, int i, int j, int iMax, int jMax) {
    return (i >= 1 && i <= iMax && j >= 1 && j <= jMax) ? T(0) + r0 / aP : T(0);
}

// Where the last Jacobi iterate's delta of one momentum system (c = 0: U,
// 1: V) of a single-precision slot is read after the fixed-budget sweeps, as
// k_mom_finish reads it: after one sweep (the assembly's) it is the implicit
// r0 / aP, else the delta buffer of the sweep count's parity. A system still
// marked running made every sweep of its budget. Load x and y at a node;
// delta = one ? delta1(x, y, ...) : x. With a stored delta both bases are the
// delta buffer, so the second load hits the line of the first.
struct DeltaSrc {
    gp<const float> x, y;
    bool one;
};
__device__ __forceinline__ DeltaSrc deltaSrc(const SlotDesc& d, const SlotState& st, int c) {
    void* const* arr = c == 0 ? d.mu : d.mv;
    const int it = st.jdone[c] ? st.jiters[c] : d.maxMomIter;
    const bool one = it == 1;
    const auto dl = G(static_cast<const float*>(arr[M_D0 + (it & 1)]));
    return {one ? G(static_cast<const float*>(arr[M_R0])) : dl, one ? G(static_cast<const float*>(arr[M_AP])) : dl, one};
}

// The arrays of one momentum system a sweep reads.
template <typename T>
struct JacSys {
    gp<const T> aP, fe, fn, r0;
};

// One Jacobi sweep of one system for a wave's R rows x 64 columns, all loads
// issued up front: delta for the rows and the rows above and below (lanes
// 0-31 also at column j0-1, lanes 32-63 at column j0+64, for the lanes at
// the wave's edges), aP and r0 of the rows, the east face fluxes of the rows
// (and of column j0-1 for lane 0's west face) and the north face fluxes of
// the rows and of the row below. The neighbour coefficients are formed from
// the fluxes as the assembly formed them (MomArr). delta is zero outside
// each system's unknowns and is taken as zero outside the node grid, so the
// product needs no masks; the coefficients of faces towards non-unknowns
// (whose fluxes the assembly leaves at 0) multiply those zeros. FromR0: the
// sweep after the assembly's, whose iterate delta_1 = r0 / aP is formed here
// for the rows, the rows above and below and the edge columns.
template <typename T, int R, bool FromR0>
__device__ __forceinline__ void jacobiWave(const Tile& tl, int M, int N, int P, int iMax, int jMax, T De, T Dn,
                                           const JacSys<T>& sy, gp<const T> src, gp<T> dst, double& v) {
    const int lane = int(threadIdx.x & 63);
    const int j = tl.j, ib = tl.i;
    const int je = lane < 32 ? tl.j0 - 1 : tl.j0 + kTileW;
    T dl[R + 2], de[R], ap[R], r0[R], fe[R], fw0[R], fn[R + 1];
    if constexpr (FromR0) {
        T rh[R + 2], ah[R + 2], re[R], ae[R];
#pragma unroll
        for (int h = 0; h < R + 2; ++h) {
            const size_t k = clampIdx(ib - 1 + h, j, M, P);
            rh[h] = sy.r0[k];
            ah[h] = sy.aP[k];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const size_t k = clampIdx(ib + r, je, M, P);
            re[r] = sy.r0[k];
            ae[r] = sy.aP[k];
        }
#pragma unroll
        for (int h = 0; h < R + 2; ++h) dl[h] = delta1(rh[h], ah[h], ib - 1 + h, j, iMax, jMax);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            de[r] = delta1(re[r], ae[r], ib + r, je, iMax, jMax);
            ap[r] = ah[r + 1];
            r0[r] = rh[r + 1];
        }
    } else {
#pragma unroll
        for (int h = 0; h < R + 2; ++h) dl[h] = src[clampIdx(ib - 1 + h, j, M, P)];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            de[r] = src[clampIdx(ib + r, je, M, P)];
            const size_t k = clampIdx(ib + r, j, M, P);
            ap[r] = sy.aP[k];
            r0[r] = sy.r0[k];
        }
        const bool jIn = j >= 0 && j <= N, jeIn = je >= 0 && je <= N;
#pragma unroll
        for (int h = 0; h < R + 2; ++h) dl[h] = (jIn && ib - 1 + h >= 0 && ib - 1 + h <= M) ? dl[h] : T(0);
#pragma unroll
        for (int r = 0; r < R; ++r) de[r] = (jeIn && ib + r <= M) ? de[r] : T(0);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        fe[r] = sy.fe[clampIdx(ib + r, j, M, P)];
        fw0[r] = sy.fe[clampIdx(ib + r, tl.j0 - 1, M, P)];
    }
#pragma unroll
    for (int h = 0; h < R + 1; ++h) fn[h] = sy.fn[clampIdx(ib - 1 + h, j, M, P)];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = ib + r;
        const T dc = dl[r + 1];
        T dW = fromWest(dc), dE = fromEast(dc), fw = fromWest(fe[r]);
        dW = lane == 0 ? de[r] : dW;
        dE = lane == 63 ? de[r] : dE;
        fw = lane == 0 ? fw0[r] : fw;
        if (i >= 1 && i <= iMax && j >= 1 && j <= jMax) {
            const T aE = De + fmax(T(0), -fe[r]);
            const T aW = De + fmax(T(0), fw);
            const T aN = Dn + fmax(T(0), -fn[r + 1]);
            const T aS = Dn + fmax(T(0), fn[r]);
            const T Ax = ap[r] * dc - aE * dE - aW * dW - aN * dl[r + 2] - aS * dl[r];
            const T res = r0[r] - Ax;
            dst[size_t(i) * P + j] = dc + res / ap[r];
            v += double(res) * double(res);
        }
    }
}

// Jacobi sweep: U and V in the same launch, each system skipped once its
// stop test has fired. U unknowns: i in [1, M-1], j in [1, N-2]; V: i in
// [1, M-2], j in [1, N-1]. After sweep s the iterate is in delta buffer
// s mod 2, except delta_1 (the assembly's), which is formed from r0 / aP.
template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_jacobi(KArgs a) {
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodeTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        const int status = st.status, doneU = st.jdone[0], doneV = st.jdone[1], jkU = st.jk[0], jkV = st.jk[1];
        if ((status != ST_RUNNING) | ((doneU != 0) & (doneV != 0))) continue;
        const int M = d.M, N = d.N, P = d.pitch;
        const T De = T(d.De), Dn = T(d.Dn);
        const int pu = jkU & 1, pv = jkV & 1;
        JacSys<T> su, sv;
        su.aP = G(static_cast<const T*>(d.mu[M_AP]));
        su.fe = G(static_cast<const T*>(d.mu[M_FE]));
        su.fn = G(static_cast<const T*>(d.mu[M_FN]));
        su.r0 = G(static_cast<const T*>(d.mu[M_R0]));
        sv.aP = G(static_cast<const T*>(d.mv[M_AP]));
        sv.fe = G(static_cast<const T*>(d.mv[M_FE]));
        sv.fn = G(static_cast<const T*>(d.mv[M_FN]));
        sv.r0 = G(static_cast<const T*>(d.mv[M_R0]));
        const auto srcU = G(static_cast<const T*>(d.mu[M_D0 + pu])), srcV = G(static_cast<const T*>(d.mv[M_D0 + pv]));
        const auto dstU = G(static_cast<T*>(d.mu[M_D0 + (pu ^ 1)])), dstV = G(static_cast<T*>(d.mv[M_D0 + (pv ^ 1)]));
        double v[2] = {0.0, 0.0};
        if (!doneU) {
            if (jkU == 1) jacobiWave<T, R, true>(tl, M, N, P, M - 1, N - 2, De, Dn, su, srcU, dstU, v[0]);
            else jacobiWave<T, R, false>(tl, M, N, P, M - 1, N - 2, De, Dn, su, srcU, dstU, v[0]);
        }
        if (!doneV) {
            if (jkV == 1) jacobiWave<T, R, true>(tl, M, N, P, M - 2, N - 1, De, Dn, sv, srcV, dstV, v[1]);
            else jacobiWave<T, R, false>(tl, M, N, P, M - 2, N - 1, De, Dn, sv, srcV, dstV, v[1]);
        }
        blockReduceTo<2, 0u>(v, a.partials + size_t(t) * kNumPartials, nred);
    }
}

// ---------------------------------------------------------------------------
// scatter with clamp + max residual (momentum_solver.cpp:318-335) fused with
// the velocity boundary conditions on (u*, v*) (boundaries.cpp:48-106).
// ---------------------------------------------------------------------------
template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_mom_finish(KArgs a) {
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodeTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        if ((st.status != ST_RUNNING) | (d.momType != 1)) continue;
        const int M = d.M, N = d.N, P = d.pitch, j = tl.j;
        const double Uin = st.inletVelocity;
        const double maxVel = 3.0 * fmax(fabs(Uin), 0.1);
        // a system still marked running made every sweep of its budget (the
        // fixed-budget step launches the last sweep without its finalisation)
        const int itU = st.jdone[0] ? st.jiters[0] : d.maxMomIter;
        const int itV = st.jdone[1] ? st.jiters[1] : d.maxMomIter;
        const auto dU = G(static_cast<const T*>(d.mu[M_D0 + (itU & 1)]));
        const auto dV = G(static_cast<const T*>(d.mv[M_D0 + (itV & 1)]));
        const auto u = G(d.u), vv = G(d.v), gam = G(d.gam);
        const auto us = G(d.us), vs = G(d.vs);
        // all loads first (clamped addresses), then the branchy selection
        double uk[R], um[R], vk[R], vm[R], ge[R];
        T duk[R], dum[R], dvk[R], dvm[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            const size_t kc = clampIdx(i, j, M, P), kw = clampIdx(i, j - 1, M, P);
            uk[r] = u[kc];
            um[r] = u[kw];
            vk[r] = vv[kc];
            vm[r] = vv[kw];
            ge[r] = gam[clampIdx(i, 1, M, P)];  // read at j = 0 only: one address per row
        }
        // delta of each system: after one sweep (the assembly's) the iterate
        // r0 / aP, else the delta buffer of the last sweep. Every load is
        // issued with the others, from bases chosen per slot: with a stored
        // delta the second base is the delta buffer again, whose second load
        // hits the line the first one fetched.
        const bool oneU = itU == 1, oneV = itV == 1;
        const auto xU = oneU ? G(static_cast<const T*>(d.mu[M_R0])) : dU;
        const auto yU = oneU ? G(static_cast<const T*>(d.mu[M_AP])) : dU;
        const auto xV = oneV ? G(static_cast<const T*>(d.mv[M_R0])) : dV;
        const auto yV = oneV ? G(static_cast<const T*>(d.mv[M_AP])) : dV;
        T xuk[R], yuk[R], xum[R], yum[R], xvk[R], yvk[R], xvm[R], yvm[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            const size_t kc = clampIdx(i, j, M, P), kw = clampIdx(i, j - 1, M, P);
            xuk[r] = xU[kc];
            yuk[r] = yU[kc];
            xum[r] = xU[kw];
            yum[r] = yU[kw];
            xvk[r] = xV[kc];
            yvk[r] = yV[kc];
            xvm[r] = xV[kw];
            yvm[r] = yV[kw];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            duk[r] = oneU ? delta1(xuk[r], yuk[r], i, j, M - 1, N - 2) : xuk[r];
            dum[r] = oneU ? delta1(xum[r], yum[r], i, j - 1, M - 1, N - 2) : xum[r];
            dvk[r] = oneV ? delta1(xvk[r], yvk[r], i, j, M - 2, N - 1) : xvk[r];
            dvm[r] = oneV ? delta1(xvm[r], yvm[r], i, j - 1, M - 2, N - 1) : xvm[r];
        }
        double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            if (i > M || j > N) continue;
            const size_t k = size_t(i) * P + j;
            // u* : (M+1) x N
            if (j <= N - 1) {
                double val;
                if (i == 0 || i == M) {
                    val = 0.0;
                } else if (j == 0) {
                    val = (ge[r] > 0.5) ? Uin : 0.0;  // gamma(i-1, 0) > 0.5
                } else if (j == N - 1) {
                    val = clampv(um[r] + double(dum[r]), maxVel);  // outlet: u(i,N-1) = u(i,N-2)
                } else {
                    val = clampv(uk[r] + double(duk[r]), maxVel);
                    v[0] = nanmax(v[0], fabs(val - uk[r]));
                }
                us[k] = val;
                if (!(val == val)) v[2] = 1.0;
            }
            // v* : M x (N+1)
            if (i <= M - 1) {
                double val;
                if (i == 0 || i == M - 1 || j == 0) {
                    val = 0.0;
                } else if (j == N) {
                    val = clampv(vm[r] + double(dvm[r]), maxVel);  // outlet: v(i,N) = v(i,N-1)
                } else {
                    val = clampv(vk[r] + double(dvk[r]), maxVel);
                    v[1] = nanmax(v[1], fabs(val - vk[r]));
                }
                vs[k] = val;
                if (!(val == val)) v[2] = 1.0;
            }
        }
        blockReduceTo<3, 7u>(v, a.partials + size_t(t) * kNumPartials, nred);
    }
}

// ---------------------------------------------------------------------------
// pressure correction: mass residual, coefficients, PCG start
// (pressure_solver.cpp:463-486, 648-653, 721-732)
// ---------------------------------------------------------------------------
template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_pres_assemble(KArgs a) {
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodeTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        if (st.status != ST_RUNNING) continue;
        const int M = d.M, N = d.N, P = d.pitch, j = tl.j;
        const int jMass = d.massScope ? N - 2 : N - 1;
        const double rho = d.rho, hx = d.hx, hy = d.hy;
        const auto us = G(d.us), vs = G(d.vs), dE = G(d.dE), dN = G(d.dN);
        const auto oCE = G(static_cast<T*>(d.pr[P_CE])), oCN = G(static_cast<T*>(d.pr[P_CN]));
        const auto oAP = G(static_cast<T*>(d.pr[P_AP])), oR = G(static_cast<T*>(d.pr[P_R]));
        const auto oX = G(static_cast<T*>(d.pr[P_X]));
        const auto bm = G(d.bm), pc = G(d.pc);
        // the FP64 imbalance is read by SOR and by refinement cycles after
        // the first only
        const bool keepB = (d.presType == 0) | (d.refineCycles > 1);
        // p' starts at 0 for SOR, which sweeps on it, and in FP64, where the
        // CG iterate is p' itself; a mixed-precision CG slot's p' is set whole
        // by the first refinement accumulation (p' = 0 + x), or, in the folded
        // fixed-budget step with one refinement cycle, not at all: the
        // correction reads x
        const bool zeroPc = (d.presType == 0) | (d.mixed == 0);
        // all loads first (clamped addresses)
        double uE[R], uW[R], vN[R], vS[R], dEe[R], dEw[R], dNn[R], dNs[R];
        // the finish's max |u* - u|, max |v* - v| and non-finite flag when
        // this launch forms u*, v* itself
        double fin[3] = {0.0, 0.0, 0.0};
        bool folded = false;
        if constexpr (sizeof(T) == 4) folded = ((a.fold & kFoldMom) != 0) & (d.momType == 1);
        if (folded) {
            // u* at the cell's east and west faces and v* at its north and
            // south faces by k_mom_finish's rules: clamp(u + delta) inside,
            // the inlet profile at j = 0, the outlet column's copy of its west
            // neighbour, 0 on the walls
            const double Uin = st.inletVelocity;
            const double maxVel = 3.0 * fmax(fabs(Uin), 0.1);
            const DeltaSrc sU = deltaSrc(d, st, 0), sV = deltaSrc(d, st, 1);
            const auto u = G(d.u), vv = G(d.v), gam = G(d.gam);
            // u at the row's node and its west neighbour; v from the row
            // below the wave's first row up (row h is row tl.i - 1 + h: the
            // south face of row r is the north face of row r - 1)
            double uk[R], um[R], vh[R + 1], ge[R];
            float xuk[R], yuk[R], xum[R], yum[R], xvh[R + 1], yvh[R + 1];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = tl.i + r;
                const size_t kc = clampIdx(i, j, M, P), kw = clampIdx(i, j - 1, M, P);
                uk[r] = u[kc];
                um[r] = u[kw];
                ge[r] = gam[clampIdx(i, 1, M, P)];  // read at j = 0 only: one address per row
                xuk[r] = sU.x[kc];
                yuk[r] = sU.y[kc];
                xum[r] = sU.x[kw];
                yum[r] = sU.y[kw];
            }
#pragma unroll
            for (int h = 0; h <= R; ++h) {
                const size_t k = clampIdx(tl.i - 1 + h, j, M, P);
                vh[h] = vv[k];
                xvh[h] = sV.x[k];
                yvh[h] = sV.y[k];
            }
            double vsh[R + 1];  // v* by k_mom_finish's interior rule, rows tl.i - 1 ..
#pragma unroll
            for (int h = 0; h <= R; ++h) {
                const float dv = sV.one ? delta1(xvh[h], yvh[h], tl.i - 1 + h, j, M - 2, N - 1) : xvh[h];
                vsh[h] = clampv(vh[h] + double(dv), maxVel);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = tl.i + r;
                const float duk = sU.one ? delta1(xuk[r], yuk[r], i, j, M - 1, N - 2) : xuk[r];
                const float dum = sU.one ? delta1(xum[r], yum[r], i, j - 1, M - 1, N - 2) : xum[r];
                const double west = clampv(um[r] + double(dum), maxVel);  // u*(i, j-1), and u*(i, N-1)
                const double here = clampv(uk[r] + double(duk), maxVel);
                uE[r] = (j == N - 1) ? west : here;
                uW[r] = (j == 1) ? ((ge[r] > 0.5) ? Uin : 0.0) : west;
                vN[r] = (i == M - 1) ? 0.0 : vsh[r + 1];
                vS[r] = (i == 1) ? 0.0 : vsh[r];
                if (!(i >= 1 && i <= M - 1 && j >= 1 && j <= N - 1)) continue;
                // each u* and v* the finish wrote is one of these faces, or 0,
                // or a copy of one of them
                if (j <= N - 2) fin[0] = nanmax(fin[0], fabs(uE[r] - uk[r]));
                if (i <= M - 2) fin[1] = nanmax(fin[1], fabs(vN[r] - vh[r + 1]));
                if (!(uE[r] == uE[r]) || !(uW[r] == uW[r]) || !(vN[r] == vN[r]) || !(vS[r] == vS[r])) fin[2] = 1.0;
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = tl.i + r;
                const size_t kc = clampIdx(i, j, M, P), kw = clampIdx(i, j - 1, M, P), ks = clampIdx(i - 1, j, M, P);
                uE[r] = us[kc];
                uW[r] = us[kw];
                vN[r] = vs[kc];
                vS[r] = vs[ks];
            }
        }
        // dE at the row and its west neighbour; dN from the row below the
        // wave's first row up (the south face of row r is row r - 1's north)
        double dNh[R + 1];
        if (d.dE32) {
            const auto fE = G(reinterpret_cast<const float*>(d.dE)), fN = G(reinterpret_cast<const float*>(d.dN));
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = tl.i + r;
                dEe[r] = fE[clampIdx(i, j, M, P)];
                dEw[r] = fE[clampIdx(i, j - 1, M, P)];
            }
#pragma unroll
            for (int h = 0; h <= R; ++h) dNh[h] = fN[clampIdx(tl.i - 1 + h, j, M, P)];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = tl.i + r;
                dEe[r] = dE[clampIdx(i, j, M, P)];
                dEw[r] = dE[clampIdx(i, j - 1, M, P)];
            }
#pragma unroll
            for (int h = 0; h <= R; ++h) dNh[h] = dN[clampIdx(tl.i - 1 + h, j, M, P)];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            dNn[r] = dNh[r + 1];
            dNs[r] = dNh[r];
        }
        double v[2] = {0.0, 0.0};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            if (!(i >= 1 && i <= M - 1 && j >= 1 && j <= N - 1)) continue;
            const size_t k = size_t(i) * P + j;
            const double mE = rho * uE[r] * hy, mW = rho * uW[r] * hy;
            const double mN = rho * vN[r] * hx, mS = rho * vS[r] * hx;
            const double b = mW - mE + mS - mN;
            if (j <= jMass) v[0] = nanmax(v[0], fabs(b));
            if (j <= N - 2) {
                const double aE = rho * dEe[r] * hy;
                const double aW = (j > 1) ? rho * dEw[r] * hy : 0.0;
                const double aN = (i < M - 1) ? rho * dNn[r] * hx : 0.0;
                const double aS = (i > 1) ? rho * dNs[r] * hx : 0.0;
                double aP = aE + aW + aN + aS;
                if (aP < 1e-20) aP = 1.0;
                // single precision: non-temporal stores
                constexpr bool nt = sizeof(T) == 4;
                stH<nt>(oCE, k, T(aE));
                stH<nt>(oCN, k, T(aN));
                stH<nt>(oAP, k, T(aP));
                const T rr = T(b);
                stH<nt>(oR, k, rr);
                stH<nt>(oX, k, T(0));
                if (keepB) bm[k] = b;
                if (zeroPc) pc[k] = 0.0;
                const T z = rr / T(aP);
                v[1] += double(rr) * double(z);
            }
        }
        if (a.fold & kFoldMom) {
            // every tile of a folded launch writes five entries (zeros for a
            // slot whose finish ran), so the finalisation reads one shape
            double v5[5] = {v[0], v[1], fin[0], fin[1], fin[2]};
            blockReduceTo<5, 29u>(v5, a.partials + size_t(t) * kNumPartials, nred);
        } else {
            blockReduceTo<2, 1u>(v, a.partials + size_t(t) * kNumPartials, nred);
        }
    }
}

// p' DOF test: i in [1, M-1], j in [1, N-2] (pressure_solver.cpp:445-455)
__device__ __forceinline__ bool isPDof(int i, int j, int M, int N) {
    return i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2;
}

// ---------------------------------------------------------------------------
// PCG iteration, part 1: p_k = z_k + beta p_{k-1} (z = r / aP, formed on the
// fly for the node and its four neighbours), q = A p_k, partial p.q
// (pressure_solver.cpp:153-159, 173-181)
// ---------------------------------------------------------------------------
// p_k at (i, j): z + beta p_{k-1} on p' unknowns, 0 elsewhere. z is the
// multigrid result when Z is given, else the Jacobi z = r / aP.
template <typename T>
__device__ __forceinline__ T pAt(gp<const T> r, gp<const T> aP, gp<const T> Z, gp<const T> pold, int i, int j,
                                 int M, int N, int P, int kk, T beta) {
    if (!isPDof(i, j, M, N)) return T(0);
    const size_t k = size_t(i) * P + j;
    const T z = Z ? Z[k] : r[k] / aP[k];
    return kk == 0 ? z : z + beta * pold[k];
}

template <typename T, bool HasZ, bool First>
__device__ __forceinline__ T pLoad(gp<const T> r, gp<const T> aP, gp<const T> Z, gp<const T> pold, uint32_t k,
                                   T beta) {
    const T z = HasZ ? ld32(Z, k) : ld32(r, k) / ld32(aP, k);
    return First ? z : z + beta * ld32(pold, k);
}

// q = A p_k for one wave's R rows x 64 C columns, C adjacent columns per
// lane (one vector access per array and row). The wave loads p_k for its
// rows and for the rows above and below them (lanes 0-31 also at column
// j0-1, lanes 32-63 at column j0+64C), then marches down the rows taking
// the neighbours across lanes from the adjacent lanes. Fold (kPcgQFold): the
// rows' CG iterate also takes the two updates before this iteration,
// x = (x + alphaA p_{k-2}) + alphaB p_{k-1}.
template <typename T, int R, int C, bool HasZ, bool First, bool Fold = false>
__device__ __forceinline__ void pcgQWave(const SlotDesc& d, const Tile& tl, int kk, T beta, double& v, T alphaA = T(0),
                                         T alphaB = T(0)) {
    const int M = d.M, N = d.N, P = d.pitch;
    const auto r = G(static_cast<const T*>(d.pr[P_R]));
    const auto aP = G(static_cast<const T*>(d.pr[P_AP]));
    const auto cE = G(static_cast<const T*>(d.pr[P_CE]));
    const auto cN = G(static_cast<const T*>(d.pr[P_CN]));
    const gp<const T> Z = HasZ ? G(static_cast<const T*>(d.pr[P_Z])) : gp<const T>(nullptr);
    const auto pold = G(static_cast<const T*>(d.pr[P_PK0 + (kk & 1)]));
    const auto pnew = G(static_cast<T*>(d.pr[P_PK0 + ((kk & 1) ^ 1)]));
    const auto Q = G(static_cast<T*>(d.pr[P_Q]));
    const int lane = int(threadIdx.x & 63);
    const int j = tl.j, ib = tl.i;
    const int je = lane < 32 ? tl.j0 - 1 : tl.j0 + kTileW * C;
    // single precision: non-temporal loads of z, p_{k-1} and the rows' own
    // aP and cE, non-temporal stores of p_k and q; cN stays cached (measured
    // over the served pool of 4 x 16 slots: 48.5 us per launch in this form,
    // 49.5 us with z and p_{k-1} cached, 52.5 us with cN non-temporal too)
    constexpr bool nt = sizeof(T) == 4;
    Vec<T, C> p[R + 2], ap[R], ce[R], cn[R + 1];
    T pe[R + 2], cee[R];
    // Fold: the rows' p_{k-2} (in the buffer p_k replaces) and x, loaded
    // with the rest; x takes the two updates as p_{k-1} arrives
    [[maybe_unused]] Vec<T, C> fx[Fold ? R : 1], fp2[Fold ? R : 1];
    [[maybe_unused]] const auto X = G(static_cast<T*>(d.pr[P_X]));
    if constexpr (Fold) {
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const uint32_t kv = clampVec32<C>(ib + rr, j, M, P);
            fx[rr] = ldvH<C, nt, T>(X, kv);
            fp2[rr] = ldvH<C, nt, T>(pnew, kv);
        }
    }
#pragma unroll
    for (int h = 0; h < R + 2; ++h) {
        const int i = ib - 1 + h;
        const uint32_t kv = clampVec32<C>(i, j, M, P);
        Vec<T, C> z;
        if (HasZ) {
            z = ldvH<C, nt>(Z, kv);
        } else {
            const Vec<T, C> rv = ldv<C>(r, kv), av = ldv<C>(aP, kv);
#pragma unroll
            for (int c = 0; c < C; ++c) z[c] = rv[c] / av[c];
        }
        if (!First) {
            const Vec<T, C> po = ldvH<C, nt>(pold, kv);
            if constexpr (Fold)
                if (h >= 1 && h <= R) {
                    // the two per-iteration updates x += alpha p, in their order
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const T x1 = fx[h - 1][c] + alphaA * fp2[h - 1][c];
                        fx[h - 1][c] = x1 + alphaB * po[c];
                    }
                }
#pragma unroll
            for (int c = 0; c < C; ++c) z[c] = z[c] + beta * po[c];
        }
        const T px = pLoad<T, HasZ, First>(r, aP, Z, pold, clampIdx32(i, je, M, P), beta);
#pragma unroll
        for (int c = 0; c < C; ++c) p[h][c] = isPDof(i, j + c, M, N) ? z[c] : T(0);
        pe[h] = isPDof(i, je, M, N) ? px : T(0);
    }
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
        const uint32_t kv = clampVec32<C>(ib + rr, j, M, P);
        ap[rr] = ldvH<C, nt>(aP, kv);
        ce[rr] = ldvH<C, nt>(cE, kv);
        cee[rr] = ld32(cE, clampIdx32(ib + rr, tl.j0 - 1, M, P));  // lane 0's west face: one address
    }
#pragma unroll
    for (int h = 0; h < R + 1; ++h) cn[h] = ldv<C>(cN, clampVec32<C>(ib - 1 + h, j, M, P));
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
        const int i = ib + rr;
        const Vec<T, C> pC = p[rr + 1];
        T pW = fromWest(pC[C - 1]), pE = fromEast(pC[0]), cEw = fromWest(ce[rr][C - 1]);
        pW = lane == 0 ? pe[rr + 1] : pW;
        cEw = lane == 0 ? cee[rr] : cEw;
        pE = lane == 63 ? pe[rr + 1] : pE;
        // CSR column order S, W, P, E, N (pressure_solver.cpp:108-118)
        Vec<T, C> q;
        bool dm[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const T west = c == 0 ? pW : pC[c > 0 ? c - 1 : 0];
            const T cw = c == 0 ? cEw : ce[rr][c > 0 ? c - 1 : 0];
            const T east = c == C - 1 ? pE : pC[c < C - 1 ? c + 1 : 0];
            q[c] = -cn[rr][c] * p[rr][c] - cw * west + ap[rr][c] * pC[c] - ce[rr][c] * east - cn[rr + 1][c] * p[rr + 2][c];
            dm[c] = isPDof(i, j + c, M, N);
        }
        const uint32_t k = uint32_t(max(i, 0)) * uint32_t(P) + uint32_t(j);
        if constexpr (Fold) stMaskedV<C, nt>(X, k, fx[rr], dm);
        stMaskedV2<C, nt>(pnew, pC, Q, q, k, dm);
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (dm[c]) v += double(pC[c]) * double(q[c]);
    }
}

// The fold form holds x and p_{k-2} of its rows as well: three waves per SIMD
// at most 168 VGPRs (without the hint it keeps 182-194, two waves).
template <typename T, int R, int V>
__global__ __launch_bounds__(kBlock, (V & kPcgQFold) ? 3 : 1) void k_pcg_q(KArgs a) {
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    constexpr int C = pColsOf(sizeof(T) == 4);
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodePTile<R, C>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        // read the gate fields together and test them without short-circuit
        // branches, so their loads overlap
        const int status = st.status, pdone = st.pdone, phase = st.stepPhase, kk = st.pk;
        if ((status != ST_RUNNING) | ((pdone | phase) != 0 && !a.force) | (kk < 0)) continue;
        const T beta = T(st.beta);
        double v[1] = {0.0};
        // the preconditioner and the first iteration are uniform per slot
        if constexpr (V == kPcgQAll) {
            if (d.precond) {
                if (kk == 0) pcgQWave<T, R, C, true, true>(d, tl, kk, beta, v[0]);
                else pcgQWave<T, R, C, true, false>(d, tl, kk, beta, v[0]);
            } else {
                if (kk == 0) pcgQWave<T, R, C, false, true>(d, tl, kk, beta, v[0]);
                else pcgQWave<T, R, C, false, false>(d, tl, kk, beta, v[0]);
            }
        } else {
            // one form (KArgs::pvar): a slot of the other preconditioner is
            // another launch's; a slot at another iteration, which the launch
            // sequence never holds, gets a non-finite p.q, so that its CG
            // fails loudly instead of taking a wrong search direction
            constexpr bool mg = (V & kPcgQMg) != 0, first = (V & kPcgQFirst) != 0;
            constexpr bool fold = (V & kPcgQFold) != 0;
            if ((d.precond != 0) != mg) continue;
            if constexpr (fold) {
                if ((kk >= 2) & ((kk & 1) == 0)) {
                    pcgQWave<T, R, C, mg, false, true>(d, tl, kk, beta, v[0], T(st.alphaPrev), T(st.alpha));
                    if (t == d.ptile0 && threadIdx.x == 0) st.nFold = kk;
                } else {
                    v[0] = __builtin_nan("");
                }
            } else if ((kk == 0) == first) {
                pcgQWave<T, R, C, mg, first>(d, tl, kk, beta, v[0]);
            } else {
                v[0] = __builtin_nan("");
            }
        }
        blockReduceTo<1, 0u>(v, a.partials + size_t(t) * kNumPartials + kPcgQPartial, nred);
    }
}

// ---------------------------------------------------------------------------
// PCG iteration, part 2: alpha = r.z / p.q, x += alpha p, r -= alpha q,
// partial r.r and r.z; stop on ||r||_2 < pTol or the iteration cap
// (pressure_solver.cpp:157-182). Every workgroup sums its slot's p.q partials
// itself in one fixed order (slotSumLean), so each gets the same alpha bit
// for bit and no finalisation launch sits between the two kernels. On
// |p.q| < 1e-20 (pressure_solver.cpp:158) every workgroup skips the slot and
// the one holding its first tile ends its CG (pdone, piters). Another
// workgroup of the same launch may read pdone before or after that write:
// it skips the slot either way, from the flag or from its own p.q.
// ---------------------------------------------------------------------------
template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_pcg_update(KArgs a) {
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    constexpr int C = pColsOf(sizeof(T) == 4);
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodePTile<R, C>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        const int status = st.status, pdone = st.pdone, phase = st.stepPhase, kk = st.pk;
        if ((status != ST_RUNNING) | (pdone != 0) | (phase != 0) | (kk < 0)) continue;
        const int M = d.M, N = d.N, P = d.pitch, j = tl.j;
        const double rz = st.rz;
        const auto pk = G(static_cast<const T*>(d.pr[P_PK0 + ((kk & 1) ^ 1)]));
        const auto x = G(static_cast<T*>(d.pr[P_X]));
        const auto r = G(static_cast<T*>(d.pr[P_R]));
        const auto Q = G(static_cast<const T*>(d.pr[P_Q]));
        const auto aP = G(static_cast<const T*>(d.pr[P_AP]));
        double v[2] = {0.0, 0.0};
        // the Jacobi r.z (r / aP) feeds beta only without the multigrid
        // preconditioner, whose beta follows the V-cycle: a multigrid slot
        // neither loads aP nor divides
        const bool jacobiZ = d.precond == 0;
        // the fixed-budget step with one refinement cycle forms x late
        // (kPcgQFold): only r changes here
        const bool lateX = (sizeof(T) == 4) & ((a.fold & kFoldPc) != 0);
        // all loads first (column vectors, clamped addresses); stores only
        // on unknowns
        // single precision: every array is touched once (non-temporal)
        constexpr bool nt = sizeof(T) == 4;
        Vec<T, C> xs[R], ps[R], rs[R], qs[R], as[R];
        if (!lateX) {
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const uint32_t k = clampVec32<C>(tl.i + rr, j, M, P);
                xs[rr] = ldvH<C, nt, T>(x, k);
                ps[rr] = ldvH<C, nt>(pk, k);
            }
        }
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const uint32_t k = clampVec32<C>(tl.i + rr, j, M, P);
            rs[rr] = ldvH<C, nt, T>(r, k);
            qs[rr] = ldvH<C, nt>(Q, k);
        }
        if (jacobiZ) {
#pragma unroll
            for (int rr = 0; rr < R; ++rr) as[rr] = ldvH<C, nt>(aP, clampVec32<C>(tl.i + rr, j, M, P));
        }
        // p.q of the slot, while the loads above are in flight
        double pq[1];
        pq[0] = slotSumLean(a.partials + kPcgQPartial, d.ptile0, d.pntiles);
        if (fabs(pq[0]) < 1e-20) {
            if (t == d.ptile0 && threadIdx.x == 0) {
                st.pdone = 1;
                st.piters = kk;
            }
            continue;
        }
        const double alphaD = rz / pq[0];
        if (t == d.ptile0 && threadIdx.x == 0) {
            st.alphaPrev = st.alpha;
            st.alpha = alphaD;
            st.nUpd = kk + 1;
        }
        const T alpha = T(alphaD);
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const int i = tl.i + rr;
            bool dm[C];
            Vec<T, C> rn;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                dm[c] = isPDof(i, j + c, M, N);
                rn[c] = rs[rr][c] - alpha * qs[rr][c];
            }
            const uint32_t k = uint32_t(max(i, 0)) * uint32_t(P) + uint32_t(j);
            if (lateX) {
                stMaskedV<C, nt>(r, k, rn, dm);
            } else {
                Vec<T, C> xn;
#pragma unroll
                for (int c = 0; c < C; ++c) xn[c] = xs[rr][c] + alpha * ps[rr][c];
                stMaskedV2<C, nt>(x, xn, r, rn, k, dm);
            }
#pragma unroll
            for (int c = 0; c < C; ++c)
                if (dm[c]) v[0] += double(rn[c]) * double(rn[c]);
            if (jacobiZ) {
#pragma unroll
                for (int c = 0; c < C; ++c)
                    if (dm[c]) v[1] += double(rn[c]) * double(rn[c] / as[rr][c]);
            }
        }
        blockReduceTo<2, 0u>(v, a.partials + size_t(t) * kNumPartials, nred);
    }
}

// ---------------------------------------------------------------------------
// Mixed mode, FP64 iterative refinement of p': p' += x (FP32 correction),
// then (if another cycle follows) r64 = b - A p' in FP64, r = r64, x = 0.
// ---------------------------------------------------------------------------
template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_refine_accum(KArgs a) {
    const int nT = *a.nTiles;
    constexpr int C = pColsOf(sizeof(T) == 4);
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodePTile<R, C>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        if (st.status != ST_RUNNING || st.stepPhase) continue;
        const int M = d.M, N = d.N, P = d.pitch, j = tl.j;
        const auto x = G(static_cast<const T*>(d.pr[P_X]));
        const auto pc = G(d.pc);
        // all loads first (column vectors, clamped addresses); the first
        // cycle accumulates onto p' = 0 (the pressure assembly does not clear
        // it)
        const bool first = st.pcycle == 0;
        Vec<T, C> xs[R];
        Vec<double, C> ps[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t k = clampVec32<C>(tl.i + r, j, M, P);
            xs[r] = ldv<C>(x, k);
#pragma unroll
            for (int c = 0; c < C; ++c) ps[r][c] = 0.0;
        }
        if (!first) {
#pragma unroll
            for (int r = 0; r < R; ++r) ps[r] = ldv<C, double>(pc, clampVec32<C>(tl.i + r, j, M, P));
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            Vec<double, C> pn;
            bool dm[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                pn[c] = ps[r][c] + double(xs[r][c]);
                dm[c] = isPDof(i, j + c, M, N);
            }
            stMaskedV<C>(pc, uint32_t(i) * uint32_t(P) + uint32_t(j), pn, dm);
        }
    }
}

template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_refine_resid(KArgs a) {
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    constexpr int C = pColsOf(sizeof(T) == 4);
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodePTile<R, C>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        if (st.status != ST_RUNNING || st.stepPhase) continue;
        const int M = d.M, N = d.N, P = d.pitch;
        const auto aP = G(static_cast<const T*>(d.pr[P_AP]));
        const auto cE = G(static_cast<const T*>(d.pr[P_CE]));
        const auto cN = G(static_cast<const T*>(d.pr[P_CN]));
        const auto pc = G(d.pc);
        double v[2] = {0.0, 0.0};
#pragma unroll
        for (int n = 0; n < C * R; ++n) {
            const int i = tl.i + n / C, j = tl.j + n % C;
            if (!isPDof(i, j, M, N)) continue;
            const size_t k = size_t(i) * P + j;
            auto pAt64 = [&](int ii, int jj, size_t kn) -> double { return isPDof(ii, jj, M, N) ? pc[kn] : 0.0; };
            const double Ap = -double(cN[k - P]) * pAt64(i - 1, j, k - P) - double(cE[k - 1]) * pAt64(i, j - 1, k - 1) +
                              double(aP[k]) * pc[k] - double(cE[k]) * pAt64(i, j + 1, k + 1) -
                              double(cN[k]) * pAt64(i + 1, j, k + P);
            const double r64 = G(d.bm)[k] - Ap;
            const T r = T(r64);
            G(static_cast<T*>(d.pr[P_R]))[k] = r;
            G(static_cast<T*>(d.pr[P_X]))[k] = T(0);
            v[0] += r64 * r64;
            v[1] += double(r) * double(r / aP[k]);
        }
        blockReduceTo<2, 0u>(v, a.partials + size_t(t) * kNumPartials, nred);
    }
}

// ---------------------------------------------------------------------------
// Red-black SOR on p' (pressure_solver.cpp:809-909, pressureSolverType 0),
// in FP64 on the p' field itself: one colour per launch; the cells of a
// colour couple only to the other colour, so each sweep is a parallel update
// of the reference's in-place one. The partial of a tile is max |change|
// over its cells of both colours (the second colour adds to the first's),
// NaN changes ignored as the reference's comparison does.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kBlock) void k_sor_sweep(KArgs a) {
    const int nT = *a.nTiles;
    __shared__ double red[1];
    int nred = 0;  // block reductions made so far (alternating buffers)
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodeTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        const SlotState& st = a.state[tl.s];
        const int status = st.status, done = st.sorDone;
        if ((status != ST_RUNNING) | (d.presType != 0) | (done != 0)) continue;
        const int M = d.M, N = d.N, P = d.pitch, j = tl.j;
        const double rho = d.rho, hx = d.hx, hy = d.hy, om = d.sorOmega;
        const auto dE = G(d.dE), dN = G(d.dN), bm = G(d.bm), pc = G(d.pc);
        const auto fE = G(reinterpret_cast<const float*>(d.dE)), fN = G(reinterpret_cast<const float*>(d.dN));
        const bool e32 = d.dE32 != 0;
        auto ldE = [&](size_t k) { return e32 ? double(fE[k]) : dE[k]; };
        auto ldN = [&](size_t k) { return e32 ? double(fN[k]) : dN[k]; };
        double mc = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tl.i + r;
            if (!(i >= 1 && i <= M - 1 && j >= 1 && j <= N - 1) || ((i + j) & 1) != a.color) continue;
            const size_t k = size_t(i) * P + j;
            const double aE = (j < N - 1) ? rho * ldE(k) * hy : 0.0;
            const double aW = (j > 1) ? rho * ldE(k - 1) * hy : 0.0;
            const double aN = (i < M - 1) ? rho * ldN(k) * hx : 0.0;
            const double aS = (i > 1) ? rho * ldN(k - P) * hx : 0.0;
            const double aP = aE + aW + aN + aS;
            if (j == N - 1 || aP < 1e-20) {
                pc[k] = 0.0;
                continue;
            }
            const double b = bm[k];  // mass imbalance of the cell (k_pres_assemble)
            const double pE = (j < N - 1) ? pc[k + 1] : 0.0;
            const double pW = (j > 1) ? pc[k - 1] : 0.0;
            const double pN = (i < M - 1) ? pc[k + P] : 0.0;
            const double pS = (i > 1) ? pc[k - P] : 0.0;
            const double pGS = (aE * pE + aW * pW + aN * pN + aS * pS + b) / aP;
            const double cur = pc[k];
            double pNew = cur + om * (pGS - cur);
            pNew = fmax(-10000.0, fmin(10000.0, pNew));
            const double ch = fabs(pNew - cur);
            if (ch > mc) mc = ch;
            pc[k] = pNew;
        }
        double v[1] = {mc};
        blockReduceTo<1, 1u>(v, red, nred);
        if (threadIdx.x == 0) {
            double* part = a.partials + size_t(t) * kNumPartials;
            double m = red[0];
            if (a.color == 1 && part[0] > m) m = part[0];
            part[0] = m;
        }
    }
}

// ---------------------------------------------------------------------------
// Aggregation-multigrid preconditioner (pressurePrecond = 1). One V-cycle:
// at each level two damped-Jacobi sweeps from zero, the residual summed over
// 2x2 aggregates into the next level's right-hand side; the coarsest level
// solved with a dense inverse; then, going up, the coarse correction (scaled)
// added to the pre-smoothed iterate and two more damped-Jacobi sweeps. Each
// level is one launch down and one up: a wave holds its 60 x R unknowns plus
// a two-deep halo in registers and recomputes the halo's sweeps, so the two
// sweeps and the residual need no grid-wide synchronisation.
// ---------------------------------------------------------------------------
struct MgTile {
    int s;       // slot
    int I0, J0;  // first unknown of this wave's part of the tile
};

template <int R>
__device__ __forceinline__ MgTile decodeMgTile(const KArgs& a, int t) {
    MgTile tl;
    tl.s = __builtin_amdgcn_readfirstlane(a.tileSlot[t]);
    const MgLevel& L = a.desc[tl.s].lv[a.level];
    const int lt = t - L.tile0;
    const int ty = lt / L.tilesX;
    const int tx = lt - ty * L.tilesX;
    tl.I0 = ty * R;
    tl.J0 = tx * kMgTileCols + int(threadIdx.x >> 6) * kMgTileW;  // this wave's first column
    return tl;
}

// multigrid work of a slot in the current pressure solve
__device__ __forceinline__ bool mgActive(const SlotDesc& d, const SlotState& st) {
    // no short-circuit: the four loads overlap
    const int pre = d.precond, status = st.status, pdone = st.pdone, phase = st.stepPhase;
    return (pre != 0) & (status == ST_RUNNING) & (pdone == 0) & (phase == 0);
}

// The state fields mgActive tests, for the per-slot kernels: loaded before
// the descriptor is staged, so both arrive in one round trip.
struct MgGate {
    int status, pdone, phase;
};
__device__ __forceinline__ MgGate mgGate(const SlotState& st) { return {st.status, st.pdone, st.stepPhase}; }
__device__ __forceinline__ bool mgActive(const SlotDesc& d, const MgGate& g) {
    return (d.precond != 0) & (g.status == ST_RUNNING) & (g.pdone == 0) & (g.phase == 0);
}

__device__ __forceinline__ size_t mgAt(const MgLevel& L, int I, int J) {
    return size_t(I + 1) * L.pitch + (J + 1);
}

// Level shape and arrays copied to registers once per tile (reading them
// through the descriptor inside the loops would reload them after every
// store, since the compiler cannot rule out aliasing).
template <typename T>
struct MgView {
    int H, W, P;
    gp<const T> aP;
    gp<const T> cE;
    gp<const T> cN;
    gp<const T> b;
    gp<const T> x;
    gp<T> xo;
    gp<T> z;
    gp<T> bo;
    __device__ __forceinline__ explicit MgView(const MgLevel& L)
        : H(L.H), W(L.W), P(L.pitch), aP(G(static_cast<const T*>(L.aP))), cE(G(static_cast<const T*>(L.cE))),
          cN(G(static_cast<const T*>(L.cN))), b(G(static_cast<const T*>(L.b))), x(G(static_cast<const T*>(L.x))),
          xo(G(static_cast<T*>(L.x))), z(G(static_cast<T*>(L.z))), bo(G(static_cast<T*>(L.b))) {}
    __device__ __forceinline__ bool in(int I, int J) const { return I >= 0 && I < H && J >= 0 && J < W; }
    // element index of (I, J) inside the level (0 <= I < H, 0 <= J < W), in
    // 32 bits so that ld32 / st32 address it from the array's scalar base
    __device__ __forceinline__ uint32_t at(int I, int J) const {
        return uint32_t(I + 1) * uint32_t(P) + uint32_t(J + 1);
    }
    // index of (I, J) clamped into the level (a valid address for any I, J)
    __device__ __forceinline__ uint32_t atc(int I, int J) const {
        return at(min(max(I, 0), H - 1), min(max(J, 0), W - 1));
    }
};

// (A x) at a node from its own and its four neighbours' values, terms in
// the CSR column order S, W, P, E, N (absent neighbours hold zero
// coefficients and values): cS = cN of the south node, cW = cE of the west
// node.
template <typename T>
__device__ __forceinline__ T ax5(T cS, T xS, T cW, T xW, T aPc, T xC, T cEc, T xE, T cNc, T xN) {
    T s = T(0);
    s += -cS * xS;
    s += -cW * xW;
    s += aPc * xC;
    s += -cEc * xE;
    s += -cNc * xN;
    return s;
}

// Galerkin coarse operator of level a.level (>= 1) from level a.level - 1:
// aP_c = sum of the aggregate's aP minus twice its internal couplings,
// cE_c / cN_c = sums of the couplings crossing to the east / north aggregate
// (only couplings between unknowns count).
template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_mg_coarsen(KArgs a) {
    const int nT = *a.nTiles;
    XCD_TILE_LOOP(t, nT) {
        const MgTile tl = decodeMgTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        const SlotState& st = a.state[tl.s];
        if (!mgActive(d, st) || a.level >= d.nLevels) continue;
        const MgView<T> C(d.lv[a.level]);
        const MgView<T> F(d.lv[a.level - 1]);
        const auto caP = G(static_cast<T*>(d.lv[a.level].aP));
        const auto ccE = G(static_cast<T*>(d.lv[a.level].cE));
        const auto ccN = G(static_cast<T*>(d.lv[a.level].cN));
        const int lane = int(threadIdx.x & 63);
        // the aggregate's fine aP, cE and cN, loaded together from clamped
        // addresses; only couplings between unknowns count
        T fa[R][4], fe[R][4], fn[R][4];
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const int I = tl.I0 + rr, J = tl.J0 + lane;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t k = F.atc(2 * I + (q >> 1), 2 * J + (q & 1));
                fa[rr][q] = ld32(F.aP, k);
                fe[rr][q] = ld32(F.cE, k);
                fn[rr][q] = ld32(F.cN, k);
            }
        }
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const int I = tl.I0 + rr, J = tl.J0 + lane;
            if (lane >= kMgTileW || I >= C.H || J >= C.W) continue;
            const int i0 = 2 * I, j0 = 2 * J;
            // q = 0: (i0, j0), 1: (i0, j0+1), 2: (i0+1, j0), 3: (i0+1, j0+1)
            auto g = [&](int q) { return F.in(i0 + (q >> 1), j0 + (q & 1)) ? double(fa[rr][q]) : 0.0; };
            auto eE = [&](int q) {
                const int fi = i0 + (q >> 1), fj = j0 + (q & 1);
                return (F.in(fi, fj) && fj + 1 < F.W) ? double(fe[rr][q]) : 0.0;
            };
            auto eN = [&](int q) {
                const int fi = i0 + (q >> 1), fj = j0 + (q & 1);
                return (F.in(fi, fj) && fi + 1 < F.H) ? double(fn[rr][q]) : 0.0;
            };
            double acc = 0.0;
            acc += g(0);
            acc += g(1);
            acc += g(2);
            acc += g(3);
            double internal = 0.0;
            internal += eE(0);
            internal += eE(2);
            internal += eN(0);
            internal += eN(1);
            const uint32_t k = C.at(I, J);
            st32(caP, k, T(acc - 2.0 * internal));
            st32(ccE, k, T(eE(1) + eE(3)));
            st32(ccN, k, T(eN(2) + eN(3)));
        }
    }
}

// Dense inverse of the coarsest operator (n <= kRows unknowns, 5-point on a
// grid of width W, compact in LDS at offset oc) by Gauss-Jordan elimination
// in double, by the whole block. Thread (row i = tid % kRows, column group
// cg = tid / kRows) keeps A[i][cg + kGroups m], m < kCols, in registers, so
// the block holds exactly kRows x kRows entries. Per pivot k the owners
// publish column k and row k through double-buffered LDS (one barrier per
// pivot); every thread then updates its entries. Entries outside n x n
// start at 0 and stay 0.
template <typename T, int kRows>
__device__ void gjInverse(const T* sAP, const T* sCE, const T* sCN, int oc, int n, int W, gp<double> out) {
    constexpr int kGroups = kBlock / kRows;
    constexpr int kCols = kRows / kGroups;
    static_assert(kRows * kRows == kBlock * kCols && kRows <= kCoarseMax, "kRows x kRows entries per block");
    const int tid = int(threadIdx.x);
    const int i = tid % kRows, cg = tid / kRows;
    __shared__ double fbuf[2][kCoarseMax], rbuf[2][kCoarseMax];
    double av[kCols];
#pragma unroll
    for (int m = 0; m < kCols; ++m) {
        const int r = i, col = cg + kGroups * m;
        double v = 0.0;
        if (r < n && col < n) {
            if (col == r) v = double(sAP[oc + r]);
            else if (col == r + 1 && (r % W) + 1 < W) v = -double(sCE[oc + r]);
            else if (col + 1 == r && (col % W) + 1 < W) v = -double(sCE[oc + col]);
            else if (col == r + W) v = -double(sCN[oc + r]);
            else if (col + W == r) v = -double(sCN[oc + col]);
        }
        av[m] = v;
    }
    for (int k = 0; k < n; ++k) {
        const int b = k & 1;
        if (cg == k % kGroups) {  // this column group holds column k
            double ck = 0.0;
#pragma unroll
            for (int m = 0; m < kCols; ++m) ck = (cg + kGroups * m == k) ? av[m] : ck;
            fbuf[b][i] = ck;
        }
        if (i == k) {
#pragma unroll
            for (int m = 0; m < kCols; ++m) rbuf[b][cg + kGroups * m] = av[m];
        }
        __syncthreads();
        // every LDS read of the pivot step issued before any use: a read
        // under the selects below would wait out its latency per column
        double piv = fbuf[b][k];
        const double fk = fbuf[b][i];
        double rb[kCols];
#pragma unroll
        for (int m = 0; m < kCols; ++m) rb[m] = rbuf[b][cg + kGroups * m];
        if (!(piv > 0.0)) piv = 1.0;
        const double ip = 1.0 / piv;
        const double fi = (i == k) ? 0.0 : fk;
#pragma unroll
        for (int m = 0; m < kCols; ++m) {
            const int j = cg + kGroups * m;
            const double rk = ((j == k) ? 1.0 : rb[m]) * ip;
            av[m] = (i == k) ? rk : ((j == k) ? 0.0 : av[m]) - fi * rk;
        }
    }
    if (i < n) {
#pragma unroll
        for (int m = 0; m < kCols; ++m) {
            const int j = cg + kGroups * m;
            if (j < n) out[size_t(i) * n + j] = av[m];
        }
    }
}

// In-LDS coarsening of the levels a.level .. nLevels-1 and the dense
// inverse of the coarsest level for one slot (see k_mg_factor).
template <typename T>
__device__ void mgFactorSlot(const KArgs& a, const SlotDesc& d) {
    constexpr int kCap = sizeof(T) == 4 ? kTailCapF32 : kTailCapF64;
    __shared__ T sAP[kCap], sCE[kCap], sCN[kCap];
    const int tid = int(threadIdx.x);
    const int L = d.nLevels;
    const int src = min(max(a.level, 1) - 1, L - 1);  // first level read from global memory
    __shared__ int off[kMaxLevels + 1];  // compact offsets of levels src .. L-1 (read after the first barrier)
    if (tid == 0) {
        off[src] = 0;
        for (int l = src; l < L; ++l) off[l + 1] = off[l] + d.lv[l].H * d.lv[l].W;
    }
    {
        const MgView<T> V(d.lv[src]);
        const int n = V.H * V.W;
        constexpr int kIt = (kCap + kBlock - 1) / kBlock;
        T la[kIt], le[kIt], ln[kIt];
#pragma unroll
        for (int it = 0; it < kIt; ++it) {  // loads first, then the LDS stores
            const int idx = min(tid + it * kBlock, n - 1);
            const int I = idx / V.W, J = idx - I * V.W;
            const size_t k = V.at(I, J);
            la[it] = V.aP[k];
            le[it] = V.cE[k];
            ln[it] = V.cN[k];
        }
#pragma unroll
        for (int it = 0; it < kIt; ++it) {
            const int idx = tid + it * kBlock;
            if (idx < n) {
                sAP[idx] = la[it];
                sCE[idx] = le[it];
                sCN[idx] = ln[it];
            }
        }
    }
    __syncthreads();
    for (int l = src + 1; l < L; ++l) {
        const int Hf = d.lv[l - 1].H, Wf = d.lv[l - 1].W, of = off[l - 1];
        const MgView<T> C(d.lv[l]);
        const auto caP = G(static_cast<T*>(d.lv[l].aP));
        const auto ccE = G(static_cast<T*>(d.lv[l].cE));
        const auto ccN = G(static_cast<T*>(d.lv[l].cN));
        for (int idx = tid; idx < C.H * C.W; idx += kBlock) {
            const int I = idx / C.W, J = idx - I * C.W;
            const int i0 = 2 * I, j0 = 2 * J;
            // q = 0: (i0, j0), 1: (i0, j0+1), 2: (i0+1, j0), 3: (i0+1, j0+1)
            auto in = [&](int q) { return i0 + (q >> 1) < Hf && j0 + (q & 1) < Wf; };
            auto at = [&](int q) { return of + (i0 + (q >> 1)) * Wf + j0 + (q & 1); };
            auto g = [&](int q) { return in(q) ? double(sAP[at(q)]) : 0.0; };
            auto eE = [&](int q) { return (in(q) && j0 + (q & 1) + 1 < Wf) ? double(sCE[at(q)]) : 0.0; };
            auto eN = [&](int q) { return (in(q) && i0 + (q >> 1) + 1 < Hf) ? double(sCN[at(q)]) : 0.0; };
            double acc = 0.0;
            acc += g(0);
            acc += g(1);
            acc += g(2);
            acc += g(3);
            double internal = 0.0;
            internal += eE(0);
            internal += eE(2);
            internal += eN(0);
            internal += eN(1);
            const T vP = T(acc - 2.0 * internal), vE = T(eE(1) + eE(3)), vN = T(eN(2) + eN(3));
            sAP[off[l] + idx] = vP;
            sCE[off[l] + idx] = vE;
            sCN[off[l] + idx] = vN;
            const size_t k = C.at(I, J);
            caP[k] = vP;
            ccE[k] = vE;
            ccN[k] = vN;
        }
        __syncthreads();
    }
    const int W = d.lv[L - 1].W, n = d.lv[L - 1].H * W, oc = off[L - 1];
    if (n <= kCoarseMax / 2) gjInverse<T, kCoarseMax / 2>(sAP, sCE, sCN, oc, n, W, G(d.coarseInv));
    else gjInverse<T, kCoarseMax>(sAP, sCE, sCN, oc, n, W, G(d.coarseInv));
}


// The coarse end of the hierarchy's setup, one workgroup per slot: the
// Galerkin operators of levels a.level .. nLevels-1 (coarse enough that the
// levels from a.level - 1 down fit in LDS), then the dense inverse of the
// coarsest operator (n = H W <= kCoarseMax unknowns) by Gauss-Jordan
// elimination in double. The arithmetic is k_mg_coarsen's and the
// elimination's, point for point. Levels above a.level are coarsened by
// the grid-wide launches before this one.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_factor(KArgs a) {
    const int s = blockIdx.x;
    __shared__ SlotDesc sd;
    const MgGate gate = mgGate(a.state[s]);
    const SlotDesc& d = descToLds(&sd, &a.desc[s]);
    if (d.ntiles == 0 || !mgActive(d, gate)) return;
    mgFactorSlot<T>(a, d);
}

// Down-sweep at level a.level (above the slot's tail): x1 = w b/aP,
// x2 = x1 + w (b - A x1)/aP, r = b - A x2; stores x2 and the aggregate sums
// of r as the next level's right-hand side. Each wave marches over its R
// rows with all loads issued up front: lanes hold columns J0-2 .. J0+61
// (outputs at lanes 2..61) and the wave loads rows I-2 .. I+R+1; values
// outside the level are zero.
template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_mg_down(KArgs a) {
    constexpr int NR = R + 4;
    const int nT = *a.nTiles;
    XCD_TILE_LOOP(t, nT) {
        const MgTile tl = decodeMgTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        const SlotState& st = a.state[tl.s];
        if (!mgActive(d, st) || a.level >= d.mgTail) continue;
        const MgView<T> L(d.lv[a.level]);
        const MgView<T> C(d.lv[a.level + 1]);
        const T om = T(d.mgOmega);
        const int lane = int(threadIdx.x & 63);
        const int J = tl.J0 - 2 + lane, ib = tl.I0;
        T b[NR], ap[NR], ce[NR], cn[NR];
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const uint32_t k = L.atc(ib - 2 + h, J);
            b[h] = ld32(L.b, k);
            ap[h] = ld32(L.aP, k);
            ce[h] = ld32(L.cE, k);
            cn[h] = ld32(L.cN, k);
        }
        T x1[NR];
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const bool in = L.in(ib - 2 + h, J);
            b[h] = in ? b[h] : T(0);
            ap[h] = in ? ap[h] : T(0);
            ce[h] = in ? ce[h] : T(0);
            cn[h] = in ? cn[h] : T(0);
            x1[h] = in ? om * (b[h] / ap[h]) : T(0);
        }
        T x2[NR];
        x2[0] = x2[NR - 1] = T(0);
#pragma unroll
        for (int h = 1; h < NR - 1; ++h) {
            const T s = ax5(cn[h - 1], x1[h - 1], fromWest(ce[h]), fromWest(x1[h]), ap[h], x1[h], ce[h],
                            fromEast(x1[h]), cn[h], x1[h + 1]);
            x2[h] = L.in(ib - 2 + h, J) ? x1[h] + om * ((b[h] - s) / ap[h]) : T(0);
        }
        const bool outLane = lane >= 2 && lane < 2 + kMgTileW;
        T r[R];
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const int h = rr + 2, I = ib + rr;
            const T s = ax5(cn[h - 1], x2[h - 1], fromWest(ce[h]), fromWest(x2[h]), ap[h], x2[h], ce[h],
                            fromEast(x2[h]), cn[h], x2[h + 1]);
            const bool in = L.in(I, J);
            r[rr] = in ? b[h] - s : T(0);
            if (in && outLane) st32(L.xo, L.at(I, J), x2[h]);
        }
        // aggregate (I, J), (I, J+1), (I+1, J), (I+1, J+1) at even lanes
#pragma unroll
        for (int m = 0; m < R / 2; ++m) {
            const T e0 = fromEast(r[2 * m]), e1 = fromEast(r[2 * m + 1]);
            const int Ic = (ib + 2 * m) >> 1, Jc = (J >> 1);
            if (outLane && !(lane & 1) && Ic < C.H && Jc < C.W) {
                T acc = T(0);
                acc += r[2 * m];
                acc += e0;
                acc += r[2 * m + 1];
                acc += e1;
                st32(C.bo, C.at(Ic, Jc), acc);
            }
        }
    }
}

// The coarse end of the V-cycle in one workgroup per slot: levels mgTail ..
// nLevels-1 live in LDS (compact H x W arrays), the sweeps of a level are
// separated by workgroup barriers instead of kernel launches, and the
// coarsest level is solved with the dense inverse. The arithmetic is the
// per-level kernels' (and the oracle's), point for point.
constexpr int kTailBlock = 1024;  // threads of the one-workgroup tail

template <typename T>
__device__ __forceinline__ T tailAx(const T* aP, const T* cE, const T* cN, const T* x, int H, int W, int I, int J) {
    const int k = I * W + J;
    T s = T(0);
    if (I > 0) s += -cN[k - W] * x[k - W];
    if (J > 0) s += -cE[k - 1] * x[k - 1];
    s += aP[k] * x[k];
    if (J + 1 < W) s += -cE[k] * x[k + 1];
    if (I + 1 < H) s += -cN[k] * x[k + W];
    return s;
}

template <typename T>
__global__ __launch_bounds__(kTailBlock) void k_mg_tail(KArgs a) {
    constexpr int kCap = sizeof(T) == 4 ? kTailCapF32 : kTailCapF64;
    __shared__ T sAP[kCap], sCE[kCap], sCN[kCap], sB[kCap], sX[kCap], sZ[kCap], sT[kCap];
    __shared__ double sc64[kCoarseMax];
    __shared__ double sInvT[kCoarseMax * kCoarseMax];  // coarsest inverse, transposed
    const int s = blockIdx.x;
    __shared__ SlotDesc sd;
    const MgGate gate = mgGate(a.state[s]);
    const SlotDesc& d = descToLds(&sd, &a.desc[s]);
    if (d.ntiles == 0 || !mgActive(d, gate)) return;
    const int L0 = d.mgTail, L = d.nLevels;
    const T om = T(d.mgOmega), sc = T(d.mgScale);
    const int tid = int(threadIdx.x);
    __shared__ int off[kMaxLevels + 1];  // compact offsets of the tail levels
    if (tid == 0) {
        off[L0] = 0;
        for (int l = L0; l < L; ++l) off[l + 1] = off[l] + d.lv[l].H * d.lv[l].W;
    }
    __syncthreads();
    {
        // every tail level's operator (and the first level's right-hand
        // side) with all loads issued before the LDS stores: element g of
        // the concatenated levels belongs to the level whose range holds it
        constexpr int kIt = (kCap + kTailBlock - 1) / kTailBlock;
        constexpr int kInvIt = kCoarseMax * kCoarseMax / kTailBlock;
        static_assert(kCoarseMax * kCoarseMax % kTailBlock == 0, "whole inverse loads per thread");
        const int total = off[L];
        const int nc = d.lv[L - 1].H * d.lv[L - 1].W, nc2 = nc * nc;
        T la[kIt], le[kIt], ln[kIt], lb[kIt];
        int lvOf[kIt];
        double li[kInvIt];  // the coarsest inverse, with the same batch of loads
#pragma unroll
        for (int it = 0; it < kInvIt; ++it) li[it] = G(d.coarseInv)[min(tid + it * kTailBlock, nc2 - 1)];
#pragma unroll
        for (int it = 0; it < kIt; ++it) {
            const int g = min(tid + it * kTailBlock, total - 1);
            int l = L0;
            while (l + 1 < L && off[l + 1] <= g) ++l;
            const MgLevel& V = d.lv[l];
            const int idx = g - off[l], I = idx / V.W, J = idx - I * V.W;
            const size_t k = mgAt(V, I, J);
            la[it] = G(static_cast<const T*>(V.aP))[k];
            le[it] = G(static_cast<const T*>(V.cE))[k];
            ln[it] = G(static_cast<const T*>(V.cN))[k];
            lb[it] = G(static_cast<const T*>(V.b))[k];
            lvOf[it] = l;
        }
#pragma unroll
        for (int it = 0; it < kIt; ++it) {
            const int g = tid + it * kTailBlock;
            if (g < total) {
                sAP[g] = la[it];
                sCE[g] = le[it];
                sCN[g] = ln[it];
                if (lvOf[it] == L0) sB[g] = lb[it];
            }
        }
#pragma unroll
        for (int it = 0; it < kInvIt; ++it) {
            const int g = tid + it * kTailBlock;
            if (g < nc2) {
                const int r = g / nc, c = g - r * nc;
                sInvT[c * nc + r] = li[it];
            }
        }
    }
    __syncthreads();
    for (int l = L0; l + 1 < L; ++l) {
        const int H = d.lv[l].H, W = d.lv[l].W, n = H * W, o = off[l];
        const T *aP = sAP + o, *cE = sCE + o, *cN = sCN + o, *b = sB + o;
        T *x1 = sT + o, *x2 = sX + o;
        for (int idx = tid; idx < n; idx += kTailBlock) x1[idx] = om * (b[idx] / aP[idx]);
        __syncthreads();
        for (int idx = tid; idx < n; idx += kTailBlock) {
            const int I = idx / W, J = idx - I * W;
            x2[idx] = x1[idx] + om * ((b[idx] - tailAx(aP, cE, cN, x1, H, W, I, J)) / aP[idx]);
        }
        __syncthreads();
        const int Hc = d.lv[l + 1].H, Wc = d.lv[l + 1].W;
        for (int idx = tid; idx < Hc * Wc; idx += kTailBlock) {
            const int Ic = idx / Wc, Jc = idx - Ic * Wc;
            T acc = T(0);
            for (int q = 0; q < 4; ++q) {
                const int I = 2 * Ic + (q >> 1), J = 2 * Jc + (q & 1);
                if (I < H && J < W) acc += b[I * W + J] - tailAx(aP, cE, cN, x2, H, W, I, J);
            }
            sB[off[l + 1] + idx] = acc;
        }
        __syncthreads();
    }
    {
        const int n = d.lv[L - 1].H * d.lv[L - 1].W, o = off[L - 1];
        if (tid < n) sc64[tid] = double(sB[o + tid]);
        __syncthreads();
        if (tid < n) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc += sInvT[k * n + tid] * sc64[k];
            sZ[o + tid] = T(acc);
        }
        __syncthreads();
    }
    for (int l = L - 2; l >= L0; --l) {
        const int H = d.lv[l].H, W = d.lv[l].W, n = H * W, o = off[l];
        const int Wc = d.lv[l + 1].W, oc = off[l + 1];
        const T *aP = sAP + o, *cE = sCE + o, *cN = sCN + o, *b = sB + o;
        T *x3 = sT + o, *x4 = sX + o, *z = sZ + o;
        for (int idx = tid; idx < n; idx += kTailBlock) {
            const int I = idx / W, J = idx - I * W;
            x3[idx] = x4[idx] + sc * sZ[oc + (I >> 1) * Wc + (J >> 1)];  // x4 still holds x2
        }
        __syncthreads();
        for (int idx = tid; idx < n; idx += kTailBlock) {
            const int I = idx / W, J = idx - I * W;
            x4[idx] = x3[idx] + om * ((b[idx] - tailAx(aP, cE, cN, x3, H, W, I, J)) / aP[idx]);
        }
        __syncthreads();
        for (int idx = tid; idx < n; idx += kTailBlock) {
            const int I = idx / W, J = idx - I * W;
            z[idx] = x4[idx] + om * ((b[idx] - tailAx(aP, cE, cN, x4, H, W, I, J)) / aP[idx]);
        }
        __syncthreads();
    }
    const MgView<T> V0(d.lv[L0]);
    for (int idx = tid; idx < V0.H * V0.W; idx += kTailBlock) {
        const int I = idx / V0.W, J = idx - I * V0.W;
        V0.z[V0.at(I, J)] = sZ[off[L0] + idx];
    }
}

// Up-sweep at level a.level (above the slot's tail): x3 = x2 + s z_c,
// x4 = x3 + w (b - A x3)/aP, z = x4 + w (b - A x4)/aP. On the finest level
// the tile also reduces r.z (b is the PCG residual there). Same wave layout
// as the down-sweep.
template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_mg_up(KArgs a) {
    constexpr int NR = R + 4;
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    XCD_TILE_LOOP(t, nT) {
        const MgTile tl = decodeMgTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        const SlotState& st = a.state[tl.s];
        // a.force (timing bursts of the finest level): also slots whose CG
        // has stopped; the recomputed z is overwritten by the next V-cycle
        // before any kernel reads it
        const bool forced = (a.force != 0) & (st.status == ST_RUNNING) & (d.precond != 0);
        if (!(mgActive(d, st) | forced) || a.level >= d.mgTail) continue;
        const MgView<T> L(d.lv[a.level]);
        const MgView<T> C(d.lv[a.level + 1]);
        const T om = T(d.mgOmega), sc = T(d.mgScale);
        const int lane = int(threadIdx.x & 63);
        const int J = tl.J0 - 2 + lane, ib = tl.I0;
        const int Jcl = min(max(J, 0), L.W - 1);
        T b[NR], ap[NR], ce[NR], cn[NR], x3[NR];
        {
            T xv[NR], zc[NR];
#pragma unroll
            for (int h = 0; h < NR; ++h) {
                const int Icl = min(max(ib - 2 + h, 0), L.H - 1);
                const uint32_t k = L.at(Icl, Jcl);
                b[h] = ld32(L.b, k);
                ap[h] = ld32(L.aP, k);
                ce[h] = ld32(L.cE, k);
                cn[h] = ld32(L.cN, k);
                xv[h] = ld32(L.x, k);
                zc[h] = ld32<T>(C.z, C.at(Icl >> 1, Jcl >> 1));
            }
#pragma unroll
            for (int h = 0; h < NR; ++h) {
                const bool in = L.in(ib - 2 + h, J);
                b[h] = in ? b[h] : T(0);
                ap[h] = in ? ap[h] : T(0);
                ce[h] = in ? ce[h] : T(0);
                cn[h] = in ? cn[h] : T(0);
                x3[h] = in ? xv[h] + sc * zc[h] : T(0);
            }
        }
        T x4[NR];
        x4[0] = x4[NR - 1] = T(0);
#pragma unroll
        for (int h = 1; h < NR - 1; ++h) {
            const T s = ax5(cn[h - 1], x3[h - 1], fromWest(ce[h]), fromWest(x3[h]), ap[h], x3[h], ce[h],
                            fromEast(x3[h]), cn[h], x3[h + 1]);
            x4[h] = L.in(ib - 2 + h, J) ? x3[h] + om * ((b[h] - s) / ap[h]) : T(0);
        }
        const bool outLane = lane >= 2 && lane < 2 + kMgTileW;
        double v[1] = {0.0};
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const int h = rr + 2, I = ib + rr;
            const T s = ax5(cn[h - 1], x4[h - 1], fromWest(ce[h]), fromWest(x4[h]), ap[h], x4[h], ce[h],
                            fromEast(x4[h]), cn[h], x4[h + 1]);
            if (outLane && L.in(I, J)) {
                const T z = x4[h] + om * ((b[h] - s) / ap[h]);
                st32(L.z, L.at(I, J), z);
                v[0] += double(b[h]) * double(z);
            }
        }
        if (a.level == 0) blockReduceTo<1, 0u>(v, a.partials + size_t(t) * kNumPartials + kMgZPartial, nred);
    }
}

// ---------------------------------------------------------------------------
// Outer-loop scalar logic of one slot: runIterations' loop test and the SER /
// line-search / CFL-ramp updates (SIMPLE.cpp:299-327, time_control.cpp:43-165).
// ---------------------------------------------------------------------------
__device__ void controller(const SlotDesc& d, SlotState& st) {
    const double error = nanmax(nanmax(st.residU, st.residV), st.residMass);
    st.error = error;
    const double mon = d.serUseMax ? error : st.residMass;
    double cfl = st.pseudoCFL;
    if (d.pseudo && d.enableSER && st.iter >= d.serMinIter) {
        const double cur = fmax(mon, 1e-15);
        if (st.serResidPrev <= 0.0) {
            st.serResidPrev = cur;
        } else {
            double ratio = st.serResidPrev / cur;
            ratio = fmax(ratio, d.serMaxDec);
            ratio = fmin(ratio, d.serMaxInc);
            double target = cfl * ratio;
            target = fmax(target, d.serCFLMin);
            target = fmin(target, d.serCFLMax);
            cfl = (1.0 - d.serSmooth) * cfl + d.serSmooth * target;
            if (!isfinite(cfl) || cfl <= 0.0) cfl = d.serCFLMin;
            st.serResidPrev = cur;
        }
    }
    if (d.pseudo && d.enableLS && d.enableSER) {
        if (st.serResidPrev <= 0.0) {
            st.lsAlpha = 1.0;
            st.lsBacktrack = 0;
        } else {
            const double ratio = mon / st.serResidPrev;
            if (ratio > d.lsTol) {
                st.lsBacktrack += 1;
                if (st.lsBacktrack <= d.lsMaxTries) {
                    st.lsAlpha = fmax(st.lsAlpha * d.lsAlphaReduce, d.lsAlphaMin);
                    cfl = fmax(cfl * d.lsAlphaReduce, d.serCFLMin);
                }
            } else {
                if (st.lsAlpha < 1.0) st.lsAlpha = fmin(st.lsAlpha * 1.1, 1.0);
                st.lsBacktrack = 0;
            }
        }
    }
    if (d.pseudo && d.enableRamp && !d.enableSER) {
        const double cur = fmax(mon, 1e-12);
        if (cur > d.rampStart) {
            cfl = d.cflInit;
        } else {
            const double ratio = d.rampStart / cur;
            const double raw = pow(ratio, d.rampExp);
            const double target = fmin(d.cflInit * raw, d.cflMax);
            cfl = (1.0 - d.rampSmooth) * cfl + d.rampSmooth * target;
            if (!isfinite(cfl) || cfl <= 0.0) cfl = d.cflInit;
        }
    }
    st.pseudoCFL = cfl;
    st.iter += 1;
    if (st.health || !isfinite(error)) {
        st.health = 1;
        st.status = ST_FAILED;
    } else if (error <= d.epsilon) {
        st.status = ST_CONVERGED;
    } else if (st.iter >= d.maxIterations) {
        st.status = ST_MAXITER;
    }
    st.inletVelocity = (d.inletRamp && d.rampSteps > 0 && st.iter < d.rampSteps)
                           ? (double(st.iter + 1) / double(d.rampSteps)) * d.targetVel
                           : d.targetVel;
}

// ---------------------------------------------------------------------------
// p += pAlpha p', pressure BCs, velocity correction, velocity BCs
// (iterations.cpp:456-500, boundaries.cpp:48-134). Its finalisation runs the
// slot's outer-loop scalar logic when the step is an outer iteration.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kBlock) void k_correct(KArgs a) {
    const int nT = *a.nTiles;
    int nred = 0;  // block reductions made so far (alternating buffers)
    XCD_TILE_LOOP(t, nT) {
        const Tile tl = decodeTile<R>(a, t);
        const SlotDesc& d = a.desc[tl.s];
        SlotState& st = a.state[tl.s];
        if (st.status != ST_RUNNING) continue;
        const int M = d.M, N = d.N, P = d.pitch, j = tl.j;
        const double Uin = st.inletVelocity;
        const double maxVel = 3.0 * fmax(fabs(Uin), 0.1);
        const double pAlpha = d.pAlpha;
        const auto gp_ = G(d.p), dE = G(d.dE), dN = G(d.dN);
        const auto fE = G(reinterpret_cast<const float*>(d.dE)), fN = G(reinterpret_cast<const float*>(d.dN));
        const auto gam = G(d.gam), u = G(d.u), vv = G(d.v);
        // p' as the single-precision CG iterate (one refinement cycle), and
        // u*, v* formed here from u, v and the last sweep's delta, in the
        // folded fixed-budget step (KArgs::fold)
        const bool pcFromX = ((a.fold & kFoldPc) != 0) & (d.presType != 0);
        const bool usFolded = ((a.fold & kFoldMom) != 0) & (d.momType == 1);
        // The outlet column of u (j = N-1) copies its west neighbour, and the
        // column j = N of v too: the lane of that neighbour writes both nodes,
        // so each lane reads p', u*, v*, dE and dN at its own node (p' also
        // east and north) and no lane reads a node another lane writes.
        double pr[R], pc0[R], pcE[R], pcN[R], usu[R], vsv[R], dEu[R], dNv[R], ge[R];
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const int i = tl.i + rr;
            const size_t kc = clampIdx(i, j, M, P);
            pr[rr] = gp_[kc];
            ge[rr] = gam[clampIdx(i, 1, M, P)];  // read at j = 0 only: one address per row
        }
        if (pcFromX) {
            const auto x = G(static_cast<const float*>(d.pr[P_X]));
            // the updates k_pcg_q has not added to x yet: p_{nFold} and, with
            // two pending, p_{nFold+1}, in the order of the per-iteration
            // updates x += alpha p
            const int nF = st.nFold, pend = st.nUpd - nF;
            const float aA = float(pend == 2 ? st.alphaPrev : st.alpha), aB = float(st.alpha);
            const auto pA = G(static_cast<const float*>(d.pr[P_PK0 + ((nF & 1) ^ 1)]));
            const auto pB = G(static_cast<const float*>(d.pr[P_PK0 + (((nF + 1) & 1) ^ 1)]));
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const int i = tl.i + rr;
                const size_t k0 = clampIdx(i, j, M, P), kE = clampIdx(i, j + 1, M, P), kN = clampIdx(i + 1, j, M, P);
                float x0 = x[k0], xE = x[kE], xN = x[kN];
                if (pend >= 1) {
                    x0 = x0 + aA * pA[k0];
                    xE = xE + aA * pA[kE];
                    xN = xN + aA * pA[kN];
                }
                if (pend == 2) {
                    x0 = x0 + aB * pB[k0];
                    xE = xE + aB * pB[kE];
                    xN = xN + aB * pB[kN];
                }
                pc0[rr] = x0;
                pcE[rr] = xE;
                pcN[rr] = xN;
            }
        } else {
            const auto pc = G(d.pc);
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const int i = tl.i + rr;
                pc0[rr] = pc[clampIdx(i, j, M, P)];
                pcE[rr] = pc[clampIdx(i, j + 1, M, P)];
                pcN[rr] = pc[clampIdx(i + 1, j, M, P)];
            }
        }
        if (usFolded) {
            // k_mom_finish's interior value clamp(u + delta); the correction
            // reads u* only inside (1 <= i <= M-1, 1 <= j <= N-2) and v* only
            // inside (1 <= i <= M-2, 1 <= j <= N-1)
            const DeltaSrc sU = deltaSrc(d, st, 0), sV = deltaSrc(d, st, 1);
            double uk[R], vk[R];
            float xu[R], yu[R], xv[R], yv[R];
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const size_t kc = clampIdx(tl.i + rr, j, M, P);
                uk[rr] = u[kc];
                vk[rr] = vv[kc];
                xu[rr] = sU.x[kc];
                yu[rr] = sU.y[kc];
                xv[rr] = sV.x[kc];
                yv[rr] = sV.y[kc];
            }
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const int i = tl.i + rr;
                const float du = sU.one ? delta1(xu[rr], yu[rr], i, j, M - 1, N - 2) : xu[rr];
                const float dv = sV.one ? delta1(xv[rr], yv[rr], i, j, M - 2, N - 1) : xv[rr];
                usu[rr] = clampv(uk[rr] + double(du), maxVel);
                vsv[rr] = clampv(vk[rr] + double(dv), maxVel);
            }
        } else {
            const auto us = G(d.us), vs = G(d.vs);
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const size_t kc = clampIdx(tl.i + rr, j, M, P);
                usu[rr] = us[kc];
                vsv[rr] = vs[kc];
            }
        }
        // dE, dN as floats or doubles (SlotDesc::dE32), a uniform choice
        if (d.dE32) {
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const size_t kc = clampIdx(tl.i + rr, j, M, P);
                dEu[rr] = fE[kc];
                dNv[rr] = fN[kc];
            }
        } else {
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const size_t kc = clampIdx(tl.i + rr, j, M, P);
                dEu[rr] = dE[kc];
                dNv[rr] = dN[kc];
            }
        }
        double bad = 0.0;
        // mixed-precision slots store p, u and v non-temporally
        const bool nt = d.mixed != 0;
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const int i = tl.i + rr;
            if (i > M || j > N) continue;
            const size_t k = size_t(i) * P + j;
            // pressure update + ghosts; each ghost has exactly one writer
            if (i >= 1 && i <= M - 1 && j >= 1 && j <= N - 1) {
                const double pn = pr[rr] + pAlpha * pc0[rr];
                stR(gp_, k, pn, nt);
                if (i == 1) gp_[k - P] = pn;      // p(0,j) = p(1,j)
                if (i == M - 1) gp_[k + P] = pn;  // p(M,j) = p(M-1,j)
                if (j == 1) gp_[k - 1] = pn;      // p(i,0) = p(i,1)
                if (j == N - 1) gp_[k + 1] = 0.0; // p(i,N) = 0
                if (!(pn == pn)) bad = 1.0;
            }
            // u: walls, inlet, inside; the outlet column from column N-2
            if (j <= N - 1) {
                if (i == 0 || i == M) {
                    u[k] = 0.0;
                } else if (j == 0) {
                    u[k] = (ge[rr] > 0.5) ? Uin : 0.0;
                } else if (j <= N - 2) {
                    const double du = clampv(dEu[rr] * (pc0[rr] - pcE[rr]), maxVel);
                    const double val = clampv(usu[rr] + du, maxVel);
                    stR(u, k, val, nt);
                    if (j == N - 2) u[k + 1] = val;  // u(i,N-1) = u(i,N-2)
                }
            }
            // v: walls, inflow column, inside; the outlet column from N-1
            if (i <= M - 1) {
                if (i == 0 || i == M - 1 || j == 0) {
                    vv[k] = 0.0;
                } else if (j <= N - 1) {
                    const double dv = clampv(dNv[rr] * (pc0[rr] - pcN[rr]), maxVel);
                    const double val = clampv(vsv[rr] + dv, maxVel);
                    stR(vv, k, val, nt);
                    if (j == N - 1) vv[k + 1] = val;  // v(i,N) = v(i,N-1)
                }
            }
        }
        double vb[1] = {bad};
        blockReduceTo<1, 1u>(vb, a.partials + size_t(t) * kNumPartials, nred);
    }
}

// ---------------------------------------------------------------------------
// samplePlaneAtX (postprocessing.cpp:53-231) by one workgroup; thread 0
// writes the result, which every thread may read after the call.
// ---------------------------------------------------------------------------
__device__ void planeBlock(const SlotDesc& d, double xq, double gammaMin, PlaneOut* result) {
    const PlaneReq rq = {0, xq, gammaMin};
    const int M = d.M, N = d.N, P = d.pitch;
    const double hx = d.hx, hy = d.hy, rho = d.rho;
    const double xMin = 0.5 * hx, xMax = (double(N) - 0.5) * hx;
    const double xc = fmax(xMin, fmin(rq.x, xMax));
    const double jf = xc / hx - 0.5;
    int cl = int(floor(jf));
    cl = max(0, min(cl, N - 2));
    const int cr = cl + 1;
    const double xl = (double(cl) + 0.5) * hx, xr = (double(cr) + 0.5) * hx;
    double tt = (fabs(xr - xl) < 1e-12) ? 0.0 : (xc - xl) / (xr - xl);
    tt = fmax(0.0, fmin(1.0, tt));
    const int pl = min(cl + 1, N), pr = min(cr + 1, N);
    const int ul = min(cl, N - 1), ur = min(cr, N - 1);
    const int vl = min(cl, N), vr = min(cr, N);
    double acc[5] = {0, 0, 0, 0, 0};  // staticArea, dynArea, totalWeighted, area, flux
    for (int i = 1 + int(threadIdx.x); i < M; i += kBlock) {
        // gamma(i-1, c) lives at gam(i, c+1)
        if (G(d.gam)[size_t(i) * P + cl + 1] < rq.gammaMin || G(d.gam)[size_t(i) * P + cr + 1] < rq.gammaMin) continue;
        const double fa = hy;
        acc[3] += fa;
        const double sp = (1.0 - tt) * G(d.p)[size_t(i) * P + pl] + tt * G(d.p)[size_t(i) * P + pr];
        const double un = (1.0 - tt) * G(d.u)[size_t(i) * P + ul] + tt * G(d.u)[size_t(i) * P + ur];
        const int vt = min(i - 1, M - 1), vb = min(i, M - 1);
        const double vT = (1.0 - tt) * G(d.v)[size_t(vt) * P + vl] + tt * G(d.v)[size_t(vt) * P + vr];
        const double vB = (1.0 - tt) * G(d.v)[size_t(vb) * P + vl] + tt * G(d.v)[size_t(vb) * P + vr];
        const double vtan = 0.5 * (vT + vB);
        const double vm = sqrt(un * un + vtan * vtan);
        const double dyn = 0.5 * rho * vm * vm;
        acc[0] += sp * fa;
        acc[1] += dyn * fa;
        const double fl = fabs(rho * un * fa);
        if (fl > 0.0) {
            acc[4] += fl;
            acc[2] += (sp + dyn) * fl;
        }
    }
    __shared__ double res[5];
    double a4[4] = {acc[0], acc[1], acc[2], acc[3]};
    blockReduceTo<4, 0u>(a4, res);
    __syncthreads();
    double a1[1] = {acc[4]};
    __shared__ double res2[1];
    blockReduceTo<1, 0u>(a1, res2);
    __syncthreads();
    if (threadIdx.x == 0) {
        PlaneOut o;
        const double area = res[3], fs = res2[0];
        o.flowArea = area;
        o.massFlux = fs;
        o.avgStatic = 0.0;
        o.avgDynamic = 0.0;
        o.avgTotal = 0.0;
        o.valid = 0;
        if (area > 0.0) {
            o.avgStatic = res[0] / area;
            o.avgDynamic = res[1] / area;
            o.valid = 1;
        }
        if (fs > 0.0) o.avgTotal = res[2] / fs;
        else if (area > 0.0) o.avgTotal = o.avgStatic + o.avgDynamic;
        *result = o;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_plane(const SlotDesc* desc, const PlaneReq* req, PlaneOut* out) {
    const PlaneReq rq = req[blockIdx.x];
    __shared__ PlaneOut o;
    planeBlock(desc[rq.slot], rq.x, rq.gammaMin, &o);
    if (threadIdx.x == 0) out[blockIdx.x] = o;
}

// ---------------------------------------------------------------------------
// Pressure-drop window exit (SIMPLE.cpp:338-409), one workgroup per slot,
// after the controller: the core-plane total-pressure drop of the iteration
// just completed enters the slope buffer and, from iteration rampSteps on,
// the moving window; a full window whose standard deviation and latest
// deviation from the mean are below dpPercent % of max(|mean|, 1) (and
// whose slope passes the gate) ends the run as converged.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_dp_monitor(KArgs a) {
    const int s = blockIdx.x;
    __shared__ SlotDesc sd;
    const SlotDesc& d = descToLds(&sd, &a.desc[s]);
    SlotState& st = a.state[s];
    const bool gate = d.slopeGate && d.slopeWindow > 0;
    if (d.ntiles == 0 || st.status != ST_RUNNING || (!d.dpConv && !gate)) return;
    __shared__ PlaneOut po[2];
    __shared__ int check;
    __shared__ double stat[2];
    planeBlock(d, d.xCoreIn, 1.0 - 1e-6, &po[0]);
    planeBlock(d, d.xCoreOut, 1.0 - 1e-6, &po[1]);
    const int it = st.iter - 1;  // index of the iteration just completed
    const int W = d.dpWindow, SW = d.slopeWindow;
    if (threadIdx.x == 0) {
        const double dp = po[0].avgTotal - po[1].avgTotal;
        st.dpLast = dp;
        if (gate) {
            if (st.slopeCount < SW) {
                G(d.dpSlope)[(st.slopeHead + st.slopeCount) % SW] = dp;
                st.slopeCount += 1;
            } else {
                G(d.dpSlope)[st.slopeHead] = dp;  // drop the oldest
                st.slopeHead = (st.slopeHead + 1) % SW;
            }
        }
        check = 0;
        if (d.dpConv && it >= d.rampSteps) {
            G(d.dpHist)[st.dpIdx] = dp;
            st.dpIdx = (st.dpIdx + 1) % W;
            if (st.dpIdx == 0) st.dpFilled = 1;
            check = st.dpFilled;
        }
    }
    __syncthreads();
    if (!check) return;
    double v[1] = {0.0};
    for (int k = int(threadIdx.x); k < W; k += kBlock) v[0] += G(d.dpHist)[k];
    blockReduceTo<1, 0u>(v, &stat[0]);
    __syncthreads();
    const double mean = stat[0] / double(W);
    double q[1] = {0.0};
    for (int k = int(threadIdx.x); k < W; k += kBlock) {
        const double df = G(d.dpHist)[k] - mean;
        q[0] += df * df;
    }
    blockReduceTo<1, 0u>(q, &stat[1]);
    __syncthreads();
    bool slopeOK = true;
    if (gate && st.slopeCount >= SW) {
        // least-squares slope over the buffer, oldest first at x = 0
        __shared__ double ls[4];
        double w4[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = int(threadIdx.x); k < st.slopeCount; k += kBlock) {
            const double x = double(k), y = G(d.dpSlope)[(st.slopeHead + k) % SW];
            w4[0] += x;
            w4[1] += y;
            w4[2] += x * x;
            w4[3] += x * y;
        }
        blockReduceTo<4, 0u>(w4, ls);
        __syncthreads();
        const double n = double(st.slopeCount);
        const double den = n * ls[2] - ls[0] * ls[0];
        const double slope = fabs(den) < 1e-12 ? 0.0 : (n * ls[3] - ls[0] * ls[1]) / den;
        slopeOK = fabs(slope) <= d.slopeMaxDelta / double(SW);
    }
    if (threadIdx.x == 0) {
        const double sd = sqrt(stat[1] / double(W));
        const double den = fmax(fabs(mean), 1.0);
        const double latest = G(d.dpHist)[(st.dpIdx - 1 + W) % W];
        if (sd / den * 100.0 < d.dpPercent && fabs(latest - mean) / den * 100.0 < d.dpPercent && slopeOK) {
            st.dpConverged = 1;
            st.status = ST_CONVERGED;
        }
    }
}

// ---------------------------------------------------------------------------
// Per-slot finalisation of a work kernel: one workgroup per slot sums the
// slot's tile partials in tile order (deterministic) and updates the slot's
// scalars. Launched right after the work kernel of `phase`; it applies the
// same skip rule as that kernel, so it only reads partials written by it.
// ---------------------------------------------------------------------------
// The PCG update's stop test and cap (pressure_solver.cpp:151, 167-178), and for
// the Jacobi preconditioner beta. Called by the whole block.
__device__ __forceinline__ void pcgUpdateFinalize(const KArgs& a, const SlotDesc& d, SlotState& st) {
    double tot[2];
    slotReduce<2, 0u>(a.partials, d.ptile0, d.pntiles, tot);
    if (threadIdx.x == 0) {
        st.rr = tot[0];
        if (sqrt(tot[0]) < d.pTol) {
            st.pdone = 1;
            st.piters = st.pk;
        } else {
            if (!d.precond) {  // multigrid: beta follows the V-cycle (PH_MG_Z)
                st.beta = tot[1] / st.rz;
                st.rz = tot[1];
            }
            st.pk += 1;
            if (st.pk >= d.maxPresIter) {
                st.pdone = 1;
                st.piters = d.maxPresIter;
            }
        }
    }
}

// r.z of the V-cycle just applied: the start of a (re)started CG (pk == 0)
// or the next search direction's beta. Called by the whole block.
__device__ __forceinline__ void mgZFinalize(const KArgs& a, const SlotDesc& d, SlotState& st) {
    double tot[1];
    slotReduce<1, 0u>(a.partials + kMgZPartial, d.lv[0].tile0, d.lv[0].ntiles, tot);
    if (threadIdx.x == 0) {
        if (st.pk == 0) {
            st.beta = 0.0;
        } else {
            st.beta = tot[0] / st.rz;
        }
        st.rz = tot[0];
    }
}

// The momentum finish's scalars (thread 0): the residuals and the
// non-finite flag of its partials; with `budgetEnds` (implicit momentum) a
// system still marked running made the budget's last sweep unfinalised.
__device__ __forceinline__ void momFinishScalars(const SlotDesc& d, SlotState& st, double residU, double residV,
                                                 double flag, bool budgetEnds) {
    st.residU = residU;
    st.residV = residV;
    if (flag != 0.0) st.health = 1;
    for (int c = 0; c < 2; ++c)
        if (budgetEnds && !st.jdone[c]) {
            st.jdone[c] = 1;
            st.jiters[c] = d.maxMomIter;
        }
    st.itU = st.jiters[0];
    st.itV = st.jiters[1];
}

__global__ __launch_bounds__(kBlock) void k_finalize(KArgs a, int phase, int outer) {
    const int s = blockIdx.x;
    // The slot's descriptor and state are read into LDS in one round trip
    // (every load issued before the LDS stores); the state is updated there
    // by thread 0 and written back whole, so its fields cost no dependent
    // global round trips. Nothing else touches the state while this kernel
    // runs.
    __shared__ SlotDesc sd;
    __shared__ SlotState sst;
    constexpr int kWords = int(sizeof(SlotState) / sizeof(uint32_t));
    constexpr int kDescWords = int(sizeof(SlotDesc) / sizeof(uint64_t));
    constexpr int kDescIt = (kDescWords + kBlock - 1) / kBlock;
    static_assert(kWords <= kBlock, "one word per thread");
    static_assert(sizeof(SlotDesc) % sizeof(uint64_t) == 0, "descriptor size");
    const auto gw = G(reinterpret_cast<uint32_t*>(&a.state[s]));
    uint32_t* lw = reinterpret_cast<uint32_t*>(&sst);
    {
        const int tid = int(threadIdx.x);
        const auto gs = G(reinterpret_cast<const uint64_t*>(&a.desc[s]));
        uint64_t* ld = reinterpret_cast<uint64_t*>(&sd);
        uint64_t dv[kDescIt];
#pragma unroll
        for (int it = 0; it < kDescIt; ++it) dv[it] = gs[min(tid + it * kBlock, kDescWords - 1)];
        const uint32_t sv = gw[min(tid, kWords - 1)];
#pragma unroll
        for (int it = 0; it < kDescIt; ++it)
            if (tid + it * kBlock < kDescWords) ld[tid + it * kBlock] = dv[it];
        if (tid < kWords) lw[tid] = sv;
    }
    __syncthreads();
    const SlotDesc& d = sd;
    const int ntiles = d.ntiles;
    SlotState& st = sst;
    if (ntiles == 0 || st.status != ST_RUNNING) return;
    switch (phase) {
        case PH_JAC:
            if (st.jdone[0] && st.jdone[1]) return;
            break;
        case PH_MG_Z:
            if (!d.precond) return;
            [[fallthrough]];
        case PH_PCG_U:
        case PH_PCGU_MGZ:
            if (st.pdone || st.stepPhase) return;
            break;
        case PH_REF_ACC:
        case PH_REF_RES:
        case PH_PCGU_REFACC:
            if (st.stepPhase) return;
            break;
        default:
            break;
    }
    switch (phase) {
        case PH_JAC: {
            double tot[2];
            slotReduce<2, 0u>(a.partials, d.tile0, d.ntiles, tot);
            if (threadIdx.x == 0) {
                const double tol2 = d.momTol * d.momTol;
                for (int c = 0; c < 2; ++c) {
                    if (st.jdone[c]) continue;
                    st.jk[c] += 1;
                    const double n = double(c == 0 ? d.nU : d.nV);
                    if (tot[c] < tol2 * n) {  // jacobi_solver.cpp:63
                        st.jdone[c] = 1;
                        st.jiters[c] = st.jk[c];
                    } else if (st.jk[c] >= d.maxMomIter) {
                        st.jdone[c] = 1;
                        st.jiters[c] = d.maxMomIter;
                    }
                }
            }
            break;
        }
        case PH_MOM_FIN:
        case PH_MOM_EXPL: {
            if (d.momType != (phase == PH_MOM_EXPL ? 0 : 1)) return;
            double tot[3];
            slotReduce<3, 7u>(a.partials, d.tile0, d.ntiles, tot);
            if (threadIdx.x == 0) momFinishScalars(d, st, tot[0], tot[1], tot[2], phase == PH_MOM_FIN);
            break;
        }
        case PH_SOR: {
            if (d.presType != 0 || st.sorDone) return;
            double tot[1];
            slotReduce<1, 1u>(a.partials, d.tile0, d.ntiles, tot);
            if (threadIdx.x == 0) {
                st.piters += 1;
                st.pitTotal = st.piters;
                // pressure_solver.cpp:906 (a NaN maximum never stops the sweeps)
                if (tot[0] < d.pTol || st.piters >= d.maxPresIter) st.sorDone = 1;
            }
            break;
        }
        case PH_PRES_ASM: {
            // a folded assembly also holds the finish's partials of an
            // implicit-momentum slot (kFoldFinPartial ..)
            double tot[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            const bool folded = (a.fold & kFoldMom) != 0;
            if (folded) {
                slotReduce<5, 29u>(a.partials, d.tile0, d.ntiles, tot);
            } else {
                double t2[2];
                slotReduce<2, 1u>(a.partials, d.tile0, d.ntiles, t2);
                tot[0] = t2[0];
                tot[1] = t2[1];
            }
            if (threadIdx.x == 0) {
                if (folded && d.momType == 1)
                    momFinishScalars(d, st, tot[kFoldFinPartial], tot[kFoldFinPartial + 1], tot[kFoldFinPartial + 2],
                                     true);
                st.residMass = tot[0];
                st.rz = tot[1];
                st.beta = 0.0;
                st.pk = 0;
                st.piters = 0;
                st.pitTotal = 0;
                st.pcycle = 0;
                st.nUpd = 0;
                st.nFold = 0;
                if (d.presType == 0) {
                    // SOR slot: the CG, multigrid and refinement kernels skip it
                    st.pdone = 1;
                    st.stepPhase = 1;
                    st.sorDone = (d.maxPresIter <= 0) ? 1 : 0;
                } else {
                    st.pdone = (d.maxPresIter <= 0) ? 1 : 0;
                    st.stepPhase = 0;  // refinement not finished
                }
            }
            break;
        }
        case PH_PCG_U:
        case PH_PCGU_MGZ: {
            pcgUpdateFinalize(a, d, st);
            if (phase == PH_PCG_U) break;
            // the V-cycle ran after the update whatever the stop test said;
            // its z only counts for a multigrid slot whose CG goes on
            __syncthreads();
            if (st.pdone || !d.precond) break;
            mgZFinalize(a, d, st);
            break;
        }
        case PH_PCGU_REFACC:
            // the last update of the cycle (skipped by the kernel for a slot
            // whose CG had stopped), then the accumulation's bookkeeping
            if (!st.pdone) pcgUpdateFinalize(a, d, st);
            __syncthreads();
            [[fallthrough]];
        case PH_REF_ACC: {
            if (threadIdx.x == 0) {
                st.pitTotal += st.piters;
                if (st.pcycle + 1 >= d.refineCycles) st.stepPhase = 1;  // refinement finished
            }
            break;
        }
        case PH_REF_RES: {
            double tot[2];
            slotReduce<2, 0u>(a.partials, d.ptile0, d.pntiles, tot);
            if (threadIdx.x == 0) {
                st.pcycle += 1;
                if (sqrt(tot[0]) < d.pTol) {
                    st.stepPhase = 1;
                } else {
                    st.rz = tot[1];
                    st.beta = 0.0;
                    st.pk = 0;
                    st.piters = 0;
                    st.pdone = 0;
                    st.nUpd = 0;
                    st.nFold = 0;
                }
            }
            break;
        }
        case PH_MG_Z:
            mgZFinalize(a, d, st);
            break;
        case PH_CORR: {
            double tot[1];
            slotReduce<1, 1u>(a.partials, d.tile0, d.ntiles, tot);
            if (threadIdx.x == 0) {
                if (tot[0] != 0.0) st.health = 1;
                st.itP = (d.mixed && d.presType == 1) ? st.pitTotal : st.piters;
                if (outer) controller(d, st);
            }
            break;
        }
        default:
            break;
    }
    __syncthreads();
    if (int(threadIdx.x) < kWords) gw[threadIdx.x] = lw[threadIdx.x];
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
// Multigrid level tiles: R even (whole aggregates per wave), 4 and 8 only in
// single precision (register footprint); the host uses no other.
template <typename T, int R>
constexpr bool mgRowsOk() {
    return R == 2 || ((R == 4 || R == 8) && sizeof(T) == 4);
}

template <typename T, int R>
static void launchTR(Kern k, int grid, hipStream_t s, const KArgs& a) {
    const dim3 g(grid), b(kBlock);
    switch (k) {
        case K_MOM_ASM: hipLaunchKernelGGL((k_mom_assemble<T, R>), g, b, 0, s, a); break;
        case K_JAC: hipLaunchKernelGGL((k_jacobi<T, R>), g, b, 0, s, a); break;
        case K_MOM_FIN: hipLaunchKernelGGL((k_mom_finish<T, R>), g, b, 0, s, a); break;
        case K_PRES_ASM: hipLaunchKernelGGL((k_pres_assemble<T, R>), g, b, 0, s, a); break;
        case K_PCG_Q:
            if constexpr (sizeof(T) == 4) {
                switch (a.pvar) {
                    case kPcgQMg: hipLaunchKernelGGL((k_pcg_q<T, R, kPcgQMg>), g, b, 0, s, a); break;
                    case kPcgQMg | kPcgQFirst:
                        hipLaunchKernelGGL((k_pcg_q<T, R, kPcgQMg | kPcgQFirst>), g, b, 0, s, a);
                        break;
                    case kPcgQJac: hipLaunchKernelGGL((k_pcg_q<T, R, kPcgQJac>), g, b, 0, s, a); break;
                    case kPcgQJac | kPcgQFirst:
                        hipLaunchKernelGGL((k_pcg_q<T, R, kPcgQJac | kPcgQFirst>), g, b, 0, s, a);
                        break;
                    case kPcgQMg | kPcgQFold:
                        hipLaunchKernelGGL((k_pcg_q<T, R, kPcgQMg | kPcgQFold>), g, b, 0, s, a);
                        break;
                    case kPcgQJac | kPcgQFold:
                        hipLaunchKernelGGL((k_pcg_q<T, R, kPcgQJac | kPcgQFold>), g, b, 0, s, a);
                        break;
                    default: hipLaunchKernelGGL((k_pcg_q<T, R, kPcgQAll>), g, b, 0, s, a); break;
                }
            } else {
                hipLaunchKernelGGL((k_pcg_q<T, R, kPcgQAll>), g, b, 0, s, a);
            }
            break;
        case K_PCG_U: hipLaunchKernelGGL((k_pcg_update<T, R>), g, b, 0, s, a); break;
        case K_REF_ACC: hipLaunchKernelGGL((k_refine_accum<T, R>), g, b, 0, s, a); break;
        case K_REF_RES: hipLaunchKernelGGL((k_refine_resid<T, R>), g, b, 0, s, a); break;
        case K_CORR:
        case K_CORR_OUTER: hipLaunchKernelGGL((k_correct<R>), g, b, 0, s, a); break;
        case K_MG_COARSEN:
            if constexpr (mgRowsOk<T, R>()) hipLaunchKernelGGL((k_mg_coarsen<T, R>), g, b, 0, s, a);
            break;
        case K_MG_DOWN:
            if constexpr (mgRowsOk<T, R>()) hipLaunchKernelGGL((k_mg_down<T, R>), g, b, 0, s, a);
            break;
        case K_MG_UP:
            if constexpr (mgRowsOk<T, R>()) hipLaunchKernelGGL((k_mg_up<T, R>), g, b, 0, s, a);
            break;
        case K_MG_FACTOR: hipLaunchKernelGGL((k_mg_factor<T>), g, b, 0, s, a); break;
        case K_MG_TAIL: hipLaunchKernelGGL((k_mg_tail<T>), g, dim3(kTailBlock), 0, s, a); break;
        case K_DP_MONITOR: hipLaunchKernelGGL(k_dp_monitor, g, b, 0, s, a); break;
        case K_MOM_EXPL: hipLaunchKernelGGL((k_mom_explicit<T, R>), g, b, 0, s, a); break;
        case K_MOM_EXPL_BC: hipLaunchKernelGGL((k_mom_expl_bc<R>), g, b, 0, s, a); break;
        case K_SOR: hipLaunchKernelGGL((k_sor_sweep<R>), g, b, 0, s, a); break;
    }
}

template <typename T>
static void launchT(Kern k, int rows, int grid, hipStream_t s, const KArgs& a) {
    switch (rows) {
        case 1: launchTR<T, 1>(k, grid, s, a); break;
        case 2: launchTR<T, 2>(k, grid, s, a); break;
        case 8: launchTR<T, 8>(k, grid, s, a); break;
        default: launchTR<T, 4>(k, grid, s, a); break;
    }
}

// finalisation phase of each work kernel (-1: none)
static int phaseOf(Kern k) {
    switch (k) {
        case K_MOM_ASM:  // the assembly makes the first sweep
        case K_JAC: return PH_JAC;
        case K_MOM_FIN: return PH_MOM_FIN;
        case K_PRES_ASM: return PH_PRES_ASM;
        case K_PCG_U: return PH_PCG_U;  // k_pcg_q's p.q is summed by k_pcg_update itself
        case K_REF_ACC: return PH_REF_ACC;
        case K_REF_RES: return PH_REF_RES;
        case K_CORR:
        case K_CORR_OUTER: return PH_CORR;
        case K_MG_UP: return PH_MG_Z;  // the finest level's up-sweep only
        case K_MOM_EXPL_BC: return PH_MOM_EXPL;
        case K_SOR: return PH_SOR;     // after the second colour only
        default: return -1;
    }
}

hipError_t launchKernel(Kern k, bool f32, int rows, int grid, int nslots, hipStream_t s, const KArgs& a,
                        hipEvent_t evBefore, hipEvent_t evAfter) {
    // external records: inside stream capture they become event-record nodes
    // that timestamp every replay (a plain record would only add an edge)
    if (evBefore) (void)hipEventRecordWithFlags(evBefore, s, hipEventRecordExternal);
    if (f32) launchT<float>(k, rows, grid, s, a);
    else launchT<double>(k, rows, grid, s, a);
    if (evAfter) (void)hipEventRecordWithFlags(evAfter, s, hipEventRecordExternal);
    const int ph = ((k == K_MG_UP && a.level != 0) || (k == K_SOR && a.color == 0)) ? -1 : phaseOf(k);
    if (ph >= 0)
        hipLaunchKernelGGL(k_finalize, dim3(nslots), dim3(kBlock), 0, s, a, ph, k == K_CORR_OUTER ? 1 : 0);
    return hipGetLastError();
}

hipError_t launchWorkOnly(Kern k, bool f32, int rows, int grid, hipStream_t s, const KArgs& a) {
    if (f32) launchT<float>(k, rows, grid, s, a);
    else launchT<double>(k, rows, grid, s, a);
    return hipGetLastError();
}

hipError_t launchFinalize(int phase, int nslots, hipStream_t s, const KArgs& a) {
    hipLaunchKernelGGL(k_finalize, dim3(nslots), dim3(kBlock), 0, s, a, phase, 0);
    return hipGetLastError();
}

hipError_t launchPlane(int nreq, hipStream_t s, const SlotDesc* desc, const PlaneReq* req, PlaneOut* out) {
    hipLaunchKernelGGL(k_plane, dim3(nreq), dim3(kBlock), 0, s, desc, req, out);
    return hipGetLastError();
}

}  // namespace cpdev
