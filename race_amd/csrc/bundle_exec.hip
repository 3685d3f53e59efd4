// bundle_exec.hip -- Gauss-Seidel and Kaczmarz over leaf bundles.
//
// One workgroup runs one bundle (a run of whole, mutually independent
// leaves of one colour phase).  TPR lanes share a row; each lane keeps E
// (local column, value) pairs of its row in registers for the whole launch.
// The bundle's local column space lives in LDS; the dependency waves of the
// leaves (rows of equal depth in RACE's sequential leaf order) run one after
// the other with a workgroup barrier in between.  Only the prologue touches
// HBM: the entry block, the outside columns of x, and the write-back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <type_traits>

#include "bundle.hpp"
#include "exec.hpp"
#include "exec_common.hpp"

namespace race {

namespace {

// matrix value storage: double, float, or bfloat16 bits (uint16_t).  The
// prologue keeps a value as loaded (ValBits) and converts it where it is
// used, after every load of the prologue has been issued.  A bfloat16 shift
// next to its load makes the wave wait for that load before it issues the
// next entry's loads (13 round trips in a row); the empty asm statement,
// which emits no instruction, keeps the compiler from moving the shift back
// next to the load.
template <typename VT>
struct ValBits {
    using T = VT;
};
template <>
struct ValBits<uint16_t> {
    using T = unsigned;
};
template <typename XT>
__device__ __forceinline__ XT valOf(double v) {
    return (XT)v;
}
template <typename XT>
__device__ __forceinline__ XT valOf(float v) {
    return (XT)v;
}
template <typename XT>
__device__ __forceinline__ XT valOf(unsigned v) {
    asm volatile("" : "+v"(v));
    return (XT)__uint_as_float(v << 16);
}

// Sum over the TPR lanes of a row.  The lanes of a row lie in one quad
// (TPR <= 4), so the partner's value comes through a DPP quad permutation,
// a VALU move with no LDS round trip: lane ^ 2 is quad_perm [2,3,0,1]
// (0x4E), lane ^ 1 is quad_perm [1,0,3,2] (0xB1).  The partner lanes of a
// row are always active together.  Same additions, in the same order, as a
// butterfly over shuffles.
template <int CTRL>
__device__ __forceinline__ float quadSwap(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double quadSwap(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int TPR, typename T>
__device__ __forceinline__ T laneSum(T v) {
    static_assert(TPR == 1 || TPR == 2 || TPR == 4, "the lanes of a row must lie in one quad");
    if (TPR == 4) v += quadSwap<0x4E>(v);
    if (TPR >= 2) v += quadSwap<0xB1>(v);
    return v;
}

// bundle handled by this workgroup: with xcd != 0, consecutive bundles of the
// phase go to the same XCD (workgroups are dealt round-robin over 8 XCDs), so
// the outside columns they share stay in one L2.  Returns -1 for the padding
// workgroups of that mapping.
__device__ __forceinline__ int bundleOfBlock(int nbnd, int xcd) {
    if (!xcd) return blockIdx.x;
    const int chunk = (nbnd + 7) >> 3;
    const int q = (blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
    return q < nbnd ? q : -1;
}

// the lane of a row that reads and writes its x (the first of its lanes)
__device__ __forceinline__ bool own(bool active, int lane) { return active && lane == 0; }

// Vectors of a bundle: the launch's x, b and n, or, for a batch of matrices
// (BATCH), the bundle's own (each matrix keeps its vectors).
template <bool BATCH, typename XT>
struct BundleVec {
    XT* x;
    const XT* b;
    int n;
    __device__ __forceinline__ BundleVec(XT* x0, const XT* b0, int n0, XT* const* bx, const XT* const* bb,
                                         const int* bn, int bnd) {
        if (BATCH) {
            x = bx[bnd];
            b = bb[bnd];
            n = bn[bnd];
        } else {
            x = x0;
            b = b0;
            n = n0;
        }
    }
};

// The prologue has three dependent memory levels: the bundle's offsets;
// then, all issued together, its entry block, row ids, inverse diagonals,
// wave numbers and outside-column ids; then the gathers of x and b.  Row
// and column ids are unsigned: widening them for the 64-bit address is then
// a zero high half, not a sign extension that has to wait for the loaded
// value.  A phase whose chains are single bundles (bundle0 >= 0) skips the
// chain table.
template <int BS, int TPR, int E, typename VT, typename XT, bool ONE_RHS, bool BATCH, int NX = 2>
__global__ void __launch_bounds__(BS) k_gs_bundle(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const uint16_t* __restrict__ slotWave,
    const int* __restrict__ rowId, const XT* __restrict__ rdiag, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, long zero, XT* __restrict__ x0,
    const XT* __restrict__ b0, int n0, int nrhs, const int* __restrict__ chainPtr, int chain0, int nchain,
    int bundle0, int xcd, XT* const* __restrict__ bX, const XT* const* __restrict__ bB,
    const int* __restrict__ bN) {
    extern __shared__ __align__(16) unsigned char smem[];
    XT* xl = reinterpret_cast<XT*>(smem);

    const int q = bundleOfBlock(nchain, xcd);
    if (q < 0) return;
    // a chain: one bundle, or the segments of a leaf too large for one bundle,
    // run in order (each segment reads what the previous one wrote)
    const int cLo = bundle0 >= 0 ? bundle0 + q : chainPtr[chain0 + q];
    const int cHi = bundle0 >= 0 ? cLo + 1 : chainPtr[chain0 + q + 1];
    auto runBundle = [&](const int bnd) {
        const BundleVec<BATCH, XT> vec(x0, b0, n0, bX, bB, bN, bnd);
        XT* __restrict__ x = vec.x;
        const XT* __restrict__ b = vec.b;
        const int n = vec.n;
        const int s0 = bSlot[bnd], nb = bSlot[bnd + 1] - s0;
        const long off = bOff[bnd];
        const int wb = bLen[bnd];
        const int nw = bWave[bnd + 1] - bWave[bnd];
        const int e0 = bExt[bnd], ne = bExt[bnd + 1] - e0;
        const int t = threadIdx.x, i = t / TPR, l = t % TPR;
        const bool active = i < nb;
        const long lanes = (long)nb * TPR;

        // Branch-free prologue.  An entry load past the bundle's width or of
        // a padding row reads the zero entry (local column 0, value 0) that
        // ends the entry arrays, at index `zero`.  Any other load whose lane
        // is idle reads element 0 of its array (every device array holds at
        // least one element); the lane ignores it, and an id read that way
        // is replaced by 0 before it indexes x.  Without branches the wait
        // before the gathers counts exactly the entry loads still in flight,
        // so the gathers overlap the entry block.  The ids the gathers need
        // go out first.
        const int sr = active ? s0 + i : 0;
        const unsigned row = (unsigned)rowId[sr];
        // ids of the first NX*BS outside columns, in one batch (columns
        // beyond are gathered round by round)
        unsigned ex[NX];
#pragma unroll
        for (int k = 0; k < NX; ++k) ex[k] = (unsigned)ext[t + k * BS < ne ? e0 + t + k * BS : 0];
        int li[E];
        typename ValBits<VT>::T lv[E];
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const long e = (active && u < wb) ? off + (long)u * lanes + t : zero;
            li[u] = (int)sidx[e];
            lv[u] = sval[e];
        }
        const XT rd = rdiag[sr];
        const int myWave = (int)slotWave[sr];

        // right-hand sides are independent: grid.y splits them over workgroups
        const int k0 = ONE_RHS ? 0 : (int)blockIdx.y;
        const int k1 = ONE_RHS ? 1 : (gridDim.y > 1 ? k0 + 1 : nrhs);
        const bool own = active && l == 0;
        for (int k = k0; k < k1; ++k) {
            const size_t vo = (size_t)k * n;
            // every gather in flight before the first LDS write
            XT g[NX];
#pragma unroll
            for (int q = 0; q < NX; ++q) g[q] = x[vo + (t + q * BS < ne ? ex[q] : 0u)];
            const XT gr = x[vo + (own ? row : 0u)];
            const XT br = b[vo + (own ? row : 0u)];
#pragma unroll
            for (int q = 0; q < NX; ++q)
                if (t + q * BS < ne) xl[nb + t + q * BS] = g[q];
            if (own) xl[i] = gr;
            for (int c = t + NX * BS; c < ne; c += BS) xl[nb + c] = x[vo + ext[e0 + c]];
            __syncthreads();
            for (int w = 0; w < nw; ++w) {
                if (active && myWave == w) {
                    // two partial sums: half the dependent FMA chain
                    XT sa = XT(0), sb = XT(0);
#pragma unroll
                    for (int u = 0; u + 1 < E; u += 2) {
                        sa += valOf<XT>(lv[u]) * xl[li[u]];
                        sb += valOf<XT>(lv[u + 1]) * xl[li[u + 1]];
                    }
                    if (E % 2) sa += valOf<XT>(lv[E - 1]) * xl[li[E - 1]];
                    const XT sm = laneSum<TPR>(sa + sb);
                    if (l == 0) xl[i] = (br - sm) * rd;
                }
                __syncthreads();
            }
            if (own) x[vo + row] = xl[i];
            __syncthreads();
        }
    };
    // the common case, one bundle per workgroup, as its own straight-line copy
    if (cHi == cLo + 1) {
        runBundle(cLo);
    } else {
        for (int bnd = cLo; bnd < cHi; ++bnd) {
            runBundle(bnd);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// One launch per sweep ("flow"): the chains of every phase are handed out in
// chain order (phase order) from a ticket counter, and a chain waits only for
// the earlier chains it conflicts with (chainDeps), through a completion flag
// per chain, instead of for the whole previous phase at a launch boundary.
// Tickets go out in an order in which every dependency comes first, and a
// workgroup takes a ticket only when it can run it, so the chain with the
// smallest unfinished number always has its dependencies done: the launch
// cannot deadlock, whatever the number of resident workgroups.
//
// Values of x that another workgroup may have written during the launch are
// read and written with agent-scope relaxed atomics (each XCD has its own
// L2; these accesses are coherent across them).  A chain publishes its flag
// after every wave of the workgroup has waited for its stores of x; a waiter
// reads the flag, then issues its loads of x.
struct FlowState {
    int ticket;  // next chain to hand out
    int done;    // workgroups finished in this launch
    int epoch;   // launch number: a chain's flag holds the epoch it finished in
    int abort;   // a wait exceeded its poll budget in this launch
    int aborts;  // launches that aborted (kept for the host)
    int tail;    // readiness-driven sweeps: next free slot of the ready queue
    int pad[2];
};

__device__ __forceinline__ int ldAgent(const int* p) {
    return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stAgent(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ldCoh(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void stCoh(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until chain a has finished in this launch.  After pollLimit polls
// (a broken dependency table), or once another wait has given up, the
// launch is marked aborted and the wait returns, so every workgroup ends.
__device__ __forceinline__ void waitChain(const int* flag, int a, int epoch, FlowState* fs, int pollLimit) {
    int polls = 0;
    while (ldAgent(&flag[a]) != epoch) {
        if ((++polls & 255) == 0 && (polls >= pollLimit || ldAgent(&fs->abort))) {
            stAgent(&fs->abort, 1);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Gauss-Seidel, one right-hand side, all phases of a sweep in one launch.
// The bundle body is k_gs_bundle's, with the dependency wait between the
// prologue's matrix loads and its gathers of x.
// NX register slots per lane for outside-column ids and values: NX*BS
// columns are gathered in one batch (bundles of a 27-point stencil at small
// leaves have about 900-2500 outside columns).
template <int BS, int TPR, int E, typename VT, typename XT, int NX = 6>
__global__ void __launch_bounds__(BS) k_gs_flow(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const uint16_t* __restrict__ slotWave,
    const int* __restrict__ rowId, const XT* __restrict__ rdiag, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, long zero, XT* x, const XT* __restrict__ b,
    const int* __restrict__ tLo, const int* __restrict__ tHi, int nchain, const int* __restrict__ depPtr,
    const int* __restrict__ dep, int* flag, FlowState* fs, int pollLimit) {
    extern __shared__ __align__(16) unsigned char smem[];
    XT* xl = reinterpret_cast<XT*>(smem);
    __shared__ int sTicket;
    const int t = threadIdx.x, i = t / TPR, l = t % TPR;
    const int epoch = ldAgent(&fs->epoch);
    if (t == 0) sTicket = atomicAdd(&fs->ticket, 1);
    __syncthreads();
    // the chain index is uniform: its table entries load through the scalar unit
    int cur = __builtin_amdgcn_readfirstlane(sTicket);
    // The previous chain's flag goes out at the next chain's first wait for
    // memory, which covers the previous chain's stores of x: their
    // completion overlaps the next prologue.  It is published before that
    // chain waits for anything, so no wait can depend on it.
    int pending = -1;
    auto publish = [&]() {
        if (pending >= 0) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (t == 0) stAgent(&flag[pending], epoch);
            pending = -1;
        }
    };
    while (cur < nchain) {
        int next = 0;
        const int cLo = tLo[cur], cHi = tHi[cur];
        for (int bnd = cLo; bnd < cHi; ++bnd) {
            const bool first = bnd == cLo;
            const int d0 = first ? depPtr[cur] : 0, nd = first ? depPtr[cur + 1] - d0 : 0;
            // branch-free: the dependency ids load with the prologue (every
            // device array holds at least one element)
            const int dj = dep[t < nd ? d0 + t : 0];
            const int s0 = bSlot[bnd], nb = bSlot[bnd + 1] - s0;
            const long off = bOff[bnd];
            const int wb = bLen[bnd];
            const int nw = bWave[bnd + 1] - bWave[bnd];
            const int e0 = bExt[bnd], ne = bExt[bnd + 1] - e0;
            const bool active = i < nb;
            const long lanes = (long)nb * TPR;
            // branch-free prologue (as in k_gs_bundle)
            const int sr = active ? s0 + i : 0;
            const unsigned row = (unsigned)rowId[sr];
            // ids of the first NX*BS outside columns, in one batch
            unsigned ex[NX];
#pragma unroll
            for (int k = 0; k < NX; ++k) ex[k] = (unsigned)ext[t + k * BS < ne ? e0 + t + k * BS : 0];
            int li[E];
            typename ValBits<VT>::T lv[E];
#pragma unroll
            for (int u = 0; u < E; ++u) {
                const long e = (active && u < wb) ? off + (long)u * lanes + t : zero;
                li[u] = (int)sidx[e];
                lv[u] = sval[e];
            }
            const XT rd = rdiag[sr];
            const int myWave = (int)slotWave[sr];
            const XT br = b[own(active, l) ? row : 0u];
            // Columns no earlier chain of the launch writes (bit 31 clear)
            // still hold the launch's starting value: they are gathered now,
            // with the matrix, through the caches.  So is the row's own x
            // (only this chain writes it).
            XT g[NX];
#pragma unroll
            for (int k = 0; k < NX; ++k) g[k] = x[t + k * BS < ne && !(ex[k] >> 31) ? ex[k] : 0u];
            const XT gr = x[own(active, l) ? row : 0u];
            // the next ticket is requested here, behind the prologue's loads
            // (the wave waits for it where it waits for them), and read after
            // this chain
            if (first && t == 0) next = atomicAdd(&fs->ticket, 1);
            publish();
            // the chains this one conflicts with, then the columns they wrote
            if (t < nd) waitChain(flag, dj, epoch, fs, pollLimit);
            for (int j = t + BS; j < nd; j += BS) waitChain(flag, dep[d0 + j], epoch, fs, pollLimit);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NX; ++k)
                if (t + k * BS < ne && (ex[k] >> 31)) g[k] = ldCoh(&x[ex[k] & 0x7fffffffu]);
#pragma unroll
            for (int k = 0; k < NX; ++k)
                if (t + k * BS < ne) xl[nb + t + k * BS] = g[k];
            if (own(active, l)) xl[i] = gr;
            for (int c = t + NX * BS; c < ne; c += BS) {
                const unsigned ec = (unsigned)ext[e0 + c];
                xl[nb + c] = (ec >> 31) ? ldCoh(&x[ec & 0x7fffffffu]) : x[ec];
            }
            __syncthreads();
            for (int w = 0; w < nw; ++w) {
                if (active && myWave == w) {
                    XT sa = XT(0), sb = XT(0);
#pragma unroll
                    for (int u = 0; u + 1 < E; u += 2) {
                        sa += valOf<XT>(lv[u]) * xl[li[u]];
                        sb += valOf<XT>(lv[u + 1]) * xl[li[u + 1]];
                    }
                    if (E % 2) sa += valOf<XT>(lv[E - 1]) * xl[li[E - 1]];
                    const XT sm = laneSum<TPR>(sa + sb);
                    if (l == 0) xl[i] = (br - sm) * rd;
                }
                __syncthreads();
            }
            if (own(active, l)) stCoh(&x[row], xl[i]);
            __syncthreads();
        }
        // the flag goes out once every wave's stores of x are complete
        pending = cur;
        if (t == 0) sTicket = next;
        __syncthreads();
        cur = __builtin_amdgcn_readfirstlane(sTicket);
    }
    publish();
    // the last workgroup to finish prepares the next launch
    if (t == 0 && atomicAdd(&fs->done, 1) == (int)gridDim.x - 1) {
        if (ldAgent(&fs->abort)) {
            atomicAdd(&fs->aborts, 1);
            stAgent(&fs->abort, 0);
        }
        stAgent(&fs->ticket, 0);
        stAgent(&fs->done, 0);
        stAgent(&fs->epoch, epoch + 1);
    }
}

// Loads and stores of x in a bundle body.  FLOW: the column id may carry
// bit 31 (written earlier in the one-launch sweep: read coherently), and
// every store is coherent; otherwise plain accesses.
template <bool FLOW, typename XT>
__device__ __forceinline__ XT ldX(const XT* x, size_t vo, unsigned id) {
    if constexpr (FLOW) {
        const XT* p = x + vo + (id & 0x7fffffffu);
        return (id >> 31) ? ldCoh(p) : *p;
    } else {
        return x[vo + id];
    }
}
template <bool FLOW, typename XT>
__device__ __forceinline__ void stX(XT* x, size_t vo, unsigned id, XT v) {
    if constexpr (FLOW)
        stCoh(x + vo + (id & 0x7fffffffu), v);
    else
        x[vo + id] = v;
}
// nothing to wait for (the phase launches)
struct NoWait {
    __device__ __forceinline__ void operator()() const {}
};

// Kaczmarz: every column the bundle touches lives in LDS; rows of one wave
// share no column.  The row norms stay FP64.  One bundle for the right-hand
// sides [k0, k1); wait() runs between the prologue's matrix loads and the
// gathers of x.
template <int BS, int TPR, int E, typename VT, typename XT, bool BATCH, bool FLOW, class Wait>
__device__ __forceinline__ void kaczBundle(
    const int bnd, const int k0, const int k1, XT* xl, Wait&& wait, const int* __restrict__ bSlot,
    const int* __restrict__ bWave, const long* __restrict__ bOff, const int* __restrict__ bLen,
    const int* __restrict__ bExt, const uint16_t* __restrict__ slotWave, const int* __restrict__ rowId,
    const double* __restrict__ rowNorm, const int* __restrict__ ext, const uint16_t* __restrict__ sidx,
    const VT* __restrict__ sval, long zero, XT* x0, const XT* __restrict__ b0, int n0, XT* const* __restrict__ bX,
    const XT* const* __restrict__ bB, const int* __restrict__ bN) {
    const BundleVec<BATCH, XT> vec(x0, b0, n0, bX, bB, bN, bnd);
    XT* __restrict__ x = vec.x;
    const XT* __restrict__ b = vec.b;
    const int n = vec.n;
    const int s0 = bSlot[bnd], nb = bSlot[bnd + 1] - s0;
    const long off = bOff[bnd];
    const int wb = bLen[bnd];
    const int nw = bWave[bnd + 1] - bWave[bnd];
    const int e0 = bExt[bnd], ne = bExt[bnd + 1] - e0;
    const int t = threadIdx.x, i = t / TPR, l = t % TPR;
    const bool active = i < nb;
    const long lanes = (long)nb * TPR;
    (void)l;

    // branch-free prologue, ids first (as in k_gs_bundle)
    const int sr = active ? s0 + i : 0;
    const unsigned row = (unsigned)rowId[sr];
    const unsigned ex0 = (unsigned)ext[t < ne ? e0 + t : 0];
    const unsigned ex1 = (unsigned)ext[t + BS < ne ? e0 + t + BS : 0];
    int li[E];
    typename ValBits<VT>::T lv[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const long e = (active && u < wb) ? off + (long)u * lanes + t : zero;
        li[u] = (int)sidx[e];
        lv[u] = sval[e];
    }
    const double rn = rowNorm[sr];
    const int myWave = (int)slotWave[sr];
    wait();

    for (int k = k0; k < k1; ++k) {
        const size_t vo = (size_t)k * n;
        const XT g0 = ldX<FLOW>(x, vo, t < ne ? ex0 : 0u);
        const XT g1 = ldX<FLOW>(x, vo, t + BS < ne ? ex1 : 0u);
        const XT br = b[vo + (active ? row : 0u)];
        if (t < ne) xl[t] = g0;
        if (t + BS < ne) xl[t + BS] = g1;
        for (int c = t + 2 * BS; c < ne; c += BS) xl[c] = ldX<FLOW>(x, vo, (unsigned)ext[e0 + c]);
        __syncthreads();
        for (int w = 0; w < nw; ++w) {
            if (active && myWave == w) {
                // the row's x entries, read once in one batch (no other
                // row of this wave writes them)
                XT cur[E];
#pragma unroll
                for (int u = 0; u < E; ++u) cur[u] = xl[li[u]];
                XT dot = XT(0);
#pragma unroll
                for (int u = 0; u < E; ++u) dot += valOf<XT>(lv[u]) * cur[u];
                dot = laneSum<TPR>(dot);
                const XT scale = (XT)(((double)dot - (double)br) / rn);
                // write back the entries with a nonzero value only: a
                // padding entry may name a column another row updates
#pragma unroll
                for (int u = 0; u < E; ++u) {
                    const XT a = valOf<XT>(lv[u]);
                    if (a != XT(0)) xl[li[u]] = cur[u] - scale * a;
                }
            }
            __syncthreads();
        }
        if (t < ne) stX<FLOW>(x, vo, ex0, xl[t]);
        if (t + BS < ne) stX<FLOW>(x, vo, ex1, xl[t + BS]);
        for (int c = t + 2 * BS; c < ne; c += BS) stX<FLOW>(x, vo, (unsigned)ext[e0 + c], xl[c]);
        __syncthreads();
    }
}

template <int BS, int TPR, int E, typename VT, typename XT, bool BATCH>
__global__ void __launch_bounds__(BS) k_kacz_bundle(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const uint16_t* __restrict__ slotWave,
    const int* __restrict__ rowId, const double* __restrict__ rowNorm, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, long zero, XT* __restrict__ x0,
    const XT* __restrict__ b0, int n0, int nrhs, const int* __restrict__ chainPtr, int chain0, int nchain, int xcd,
    int ldsLocal,
    XT* const* __restrict__ bX, const XT* const* __restrict__ bB, const int* __restrict__ bN) {
    extern __shared__ __align__(16) unsigned char smem[];
    XT* xl = reinterpret_cast<XT*>(smem);
    (void)ldsLocal;

    const int q = bundleOfBlock(nchain, xcd);
    if (q < 0) return;
    // a chain: one bundle, or the segments of a leaf too large for one bundle,
    // run in order (each segment reads what the previous one wrote)
    const int cLo = chainPtr[chain0 + q], cHi = chainPtr[chain0 + q + 1];
    // right-hand sides are independent: grid.y splits them over workgroups
    const int k0 = (int)blockIdx.y;
    const int k1 = gridDim.y > 1 ? k0 + 1 : nrhs;
    auto runBundle = [&](const int bnd) {
        kaczBundle<BS, TPR, E, VT, XT, BATCH, false>(bnd, k0, k1, xl, NoWait{}, bSlot, bWave, bOff, bLen, bExt,
                                                       slotWave, rowId, rowNorm, ext, sidx, sval, zero, x0, b0, n0,
                                                       bX, bB, bN);
    };
    // the common case, one bundle per workgroup, as its own straight-line copy
    if (cHi == cLo + 1) {
        runBundle(cLo);
    } else {
        for (int bnd = cLo; bnd < cHi; ++bnd) {
            runBundle(bnd);
            __syncthreads();
        }
    }
}

// Readiness-driven one-launch sweep.  A workgroup takes a queue position
// (ticket) and runs the chain stored there: the chains without dependencies
// occupy the first positions (roots), every other chain is appended by the
// workgroup that completes its last dependency.  So a workgroup never holds
// a chain whose dependencies are still running (with tickets in chain order,
// a matrix of a batch with a few long chains holds the slots of every other
// matrix's later phases).  No deadlock: a workgroup takes a position only
// after releasing the successors of its previous chain, and while positions
// wait for chains, some chain with finished dependencies has been queued.
// Per right-hand side (blockIdx.y) its own queue, arrival counters and state.
struct ReadyTables {
    const int* lo;             // chain -> first bundle
    const int* hi;             // chain -> end of its bundles
    const int* depPtr;         // chain -> its dependencies (their count)
    const int* succPtr;        // chain -> chains depending on it
    const int* succ;
    const int* roots;          // chains without dependencies
    unsigned* arrivals;        // per side: completed dependencies, summed over launches
    unsigned long long* queue; // per side: (epoch << 32) | chain, from position nroots on
    FlowState* fs;             // per side
    int nchain, nroots, pollLimit;
};

// the chain at queue position h (>= nroots) once it has been written, or
// -1 when the launch gives up (see waitChain)
__device__ __forceinline__ int waitQueue(const unsigned long long* q, int h, int epoch, FlowState* fs, int pollLimit) {
    int polls = 0;
    for (;;) {
        const unsigned long long v =
            __hip_atomic_load(const_cast<unsigned long long*>(&q[h]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((int)(v >> 32) == epoch) return (int)(unsigned)v;
        if ((++polls & 255) == 0 && (polls >= pollLimit || ldAgent(&fs->abort))) {
            stAgent(&fs->abort, 1);
            return -1;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

template <int BS, class Body>
__device__ __forceinline__ void readyLoop(const ReadyTables& r, Body&& body) {
    __shared__ int sChain;
    const int t = threadIdx.x;
    FlowState* fs = r.fs + blockIdx.y;
    unsigned* arr = r.arrivals + (size_t)blockIdx.y * r.nchain;
    unsigned long long* q = r.queue + (size_t)blockIdx.y * r.nchain;
    const int epoch = ldAgent(&fs->epoch);
    for (;;) {
        if (t == 0) {
            const int h = atomicAdd(&fs->ticket, 1);
            sChain = h < r.nroots ? r.roots[h] : h < r.nchain ? waitQueue(q, h, epoch, fs, r.pollLimit) : -1;
        }
        __syncthreads();
        const int c = __builtin_amdgcn_readfirstlane(sChain);
        if (c < 0) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const int cLo = r.lo[c], cHi = r.hi[c];
        for (int bnd = cLo; bnd < cHi; ++bnd) body(bnd, NoWait{});
        // every wave's stores of x are complete before any successor is released
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const int s0 = r.succPtr[c], ns = r.succPtr[c + 1] - s0;
        for (int j = t; j < ns; j += BS) {
            const int sc = r.succ[s0 + j];
            const unsigned need = (unsigned)epoch * (unsigned)(r.depPtr[sc + 1] - r.depPtr[sc]);
            if (atomicAdd(&arr[sc], 1u) + 1u == need) {
                const int slot = atomicAdd(&fs->tail, 1);
                __hip_atomic_store(&q[slot], ((unsigned long long)(unsigned)epoch << 32) | (unsigned)sc,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    // the last workgroup to finish prepares the next launch
    if (t == 0 && atomicAdd(&fs->done, 1) == (int)gridDim.x - 1) {
        if (ldAgent(&fs->abort)) {
            atomicAdd(&fs->aborts, 1);
            stAgent(&fs->abort, 0);
        }
        stAgent(&fs->ticket, 0);
        stAgent(&fs->tail, r.nroots);
        stAgent(&fs->done, 0);
        stAgent(&fs->epoch, epoch + 1);
    }
}

// Kaczmarz, all phases of a sweep in one readiness-driven launch; grid.y =
// right-hand sides.
template <int BS, int TPR, int E, typename VT, typename XT, bool BATCH>
__global__ void __launch_bounds__(BS) k_kacz_flow(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const uint16_t* __restrict__ slotWave,
    const int* __restrict__ rowId, const double* __restrict__ rowNorm, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, long zero, XT* x0, const XT* __restrict__ b0,
    int n0, XT* const* __restrict__ bX, const XT* const* __restrict__ bB, const int* __restrict__ bN,
    ReadyTables rt) {
    extern __shared__ __align__(16) unsigned char smem[];
    XT* xl = reinterpret_cast<XT*>(smem);
    const int k = (int)blockIdx.y;
    readyLoop<BS>(rt, [&](int bnd, auto wait) {
        kaczBundle<BS, TPR, E, VT, XT, BATCH, true>(bnd, k, k + 1, xl, wait, bSlot, bWave, bOff, bLen, bExt,
                                                      slotWave, rowId, rowNorm, ext, sidx, sval, zero, x0, b0, n0,
                                                      bX, bB, bN);
    });
}

// ---------------------------------------------------------------------------
// Variable lanes per row (matrices with a few very long rows): the row of
// slot i owns lanes [lane0(i), lane0(i+1)) of the bundle, each holding up to
// kVarE entries in registers.  Every lane stores its partial sum in LDS and
// the row's sum is read back in lane order (deterministic, no atomics), so
// every wave takes two barriers.
template <int BS, typename VT, typename XT, bool BATCH>
__global__ void __launch_bounds__(BS) k_gs_bundle_vl(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const uint16_t* __restrict__ slotWave,
    const int* __restrict__ rowId, const XT* __restrict__ diag, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, long zero, XT* __restrict__ x0,
    const XT* __restrict__ b0, int n0, int nrhs, const int* __restrict__ chainPtr, int chain0, int nchain, int xcd,
    int ldsLocal, int ldsRows, const int* __restrict__ lanePtr, const uint16_t* __restrict__ laneSlot,
    XT* const* __restrict__ bX, const XT* const* __restrict__ bB, const int* __restrict__ bN) {
    constexpr int E = kVarE;
    extern __shared__ __align__(16) unsigned char smem[];
    XT* xl = reinterpret_cast<XT*>(smem);
    XT* part = xl + ldsLocal;                               // BS partial sums
    int* rl0 = reinterpret_cast<int*>(part + BS);           // row -> first lane
    (void)ldsRows;

    const int q = bundleOfBlock(nchain, xcd);
    if (q < 0) return;
    const int cLo = chainPtr[chain0 + q], cHi = chainPtr[chain0 + q + 1];
    auto runBundle = [&](const int bnd) {
        const BundleVec<BATCH, XT> vec(x0, b0, n0, bX, bB, bN, bnd);
        XT* __restrict__ x = vec.x;
        const XT* __restrict__ b = vec.b;
        const int n = vec.n;
        const int s0 = bSlot[bnd], nb = bSlot[bnd + 1] - s0;
        const long off = bOff[bnd];
        const int wb = bLen[bnd];
        const int nw = bWave[bnd + 1] - bWave[bnd];
        const int e0 = bExt[bnd], ne = bExt[bnd + 1] - e0;
        const int lp = lanePtr[bnd], nl = lanePtr[bnd + 1] - lp;
        const int t = threadIdx.x;
        const bool active = t < nl;
        // branch-free prologue, ids first (as in k_gs_bundle); the first
        // lane of every row goes to LDS and is read after the barrier that
        // follows the gathers
        const unsigned ex0 = (unsigned)ext[t < ne ? e0 + t : 0];
        const unsigned ex1 = (unsigned)ext[t + BS < ne ? e0 + t + BS : 0];
        const int i = (int)laneSlot[active ? lp + t : 0];  // used by active lanes only
        const int iPrev = (int)laneSlot[active && t > 0 ? lp + t - 1 : 0];
        int li[E];
        typename ValBits<VT>::T lv[E];
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const long e = (active && u < wb) ? off + (long)u * nl + t : zero;
            li[u] = (int)sidx[e];
            lv[u] = sval[e];
        }
        const bool first = active && (t == 0 || iPrev != i);
        const int sr = active ? s0 + i : 0;
        const unsigned row = (unsigned)rowId[sr];
        const XT dg = diag[sr];  // 1 / a_rr
        const int myWave = (int)slotWave[sr];
        if (first) rl0[i] = t;
        if (t == 0) rl0[nb] = nl;
        int rowL0 = 0, rowL1 = 0;

        const int k0 = (int)blockIdx.y;
        const int k1 = gridDim.y > 1 ? k0 + 1 : nrhs;
        for (int k = k0; k < k1; ++k) {
            const size_t vo = (size_t)k * n;
            const XT g0 = x[vo + (t < ne ? ex0 : 0u)];
            const XT g1 = x[vo + (t + BS < ne ? ex1 : 0u)];
            const XT gr = x[vo + (first ? row : 0u)];
            const XT br = b[vo + (first ? row : 0u)];
            if (t < ne) xl[nb + t] = g0;
            if (t + BS < ne) xl[nb + t + BS] = g1;
            if (first) xl[i] = gr;
            for (int c = t + 2 * BS; c < ne; c += BS) xl[nb + c] = x[vo + ext[e0 + c]];
            __syncthreads();
            if (k == k0 && active) {
                rowL0 = rl0[i];
                rowL1 = rl0[i + 1];
            }
            for (int w = 0; w < nw; ++w) {
                if (active && myWave == w) {
                    XT sum = XT(0);
#pragma unroll
                    for (int u = 0; u < E; ++u) sum += valOf<XT>(lv[u]) * xl[li[u]];
                    part[t] = sum;
                }
                __syncthreads();
                if (first && myWave == w) {
                    XT s = XT(0);
                    for (int l = rowL0; l < rowL1; ++l) s += part[l];
                    xl[i] = (br - s) * dg;
                }
                __syncthreads();
            }
            if (first) x[vo + row] = xl[i];
            __syncthreads();
        }
    };
    if (cHi == cLo + 1) {
        runBundle(cLo);
    } else {
        for (int bnd = cLo; bnd < cHi; ++bnd) {
            runBundle(bnd);
            __syncthreads();
        }
    }
}

// Variable lanes per row, Kaczmarz (see kaczBundle).  part: BS partial
// sums, rl0: the first lane of every row of the bundle.
template <int BS, typename VT, typename XT, bool BATCH, bool FLOW, class Wait>
__device__ __forceinline__ void kaczBundleVL(
    const int bnd, const int k0, const int k1, XT* xl, XT* part, int* rl0, Wait&& wait,
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const uint16_t* __restrict__ slotWave,
    const int* __restrict__ rowId, const double* __restrict__ rowNorm, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, long zero, XT* x0, const XT* __restrict__ b0,
    int n0, const int* __restrict__ lanePtr, const uint16_t* __restrict__ laneSlot, XT* const* __restrict__ bX,
    const XT* const* __restrict__ bB, const int* __restrict__ bN) {
    constexpr int E = kVarE;
    const BundleVec<BATCH, XT> vec(x0, b0, n0, bX, bB, bN, bnd);
    XT* __restrict__ x = vec.x;
    const XT* __restrict__ b = vec.b;
    const int n = vec.n;
    const int s0 = bSlot[bnd], nb = bSlot[bnd + 1] - s0;
    const long off = bOff[bnd];
    const int wb = bLen[bnd];
    const int nw = bWave[bnd + 1] - bWave[bnd];
    const int e0 = bExt[bnd], ne = bExt[bnd + 1] - e0;
    const int lp = lanePtr[bnd], nl = lanePtr[bnd + 1] - lp;
    const int t = threadIdx.x;
    const bool active = t < nl;
    // branch-free prologue, ids first (as in k_gs_bundle_vl)
    const unsigned ex0 = (unsigned)ext[t < ne ? e0 + t : 0];
    const unsigned ex1 = (unsigned)ext[t + BS < ne ? e0 + t + BS : 0];
    const int i = (int)laneSlot[active ? lp + t : 0];  // used by active lanes only
    const int iPrev = (int)laneSlot[active && t > 0 ? lp + t - 1 : 0];
    int li[E];
    typename ValBits<VT>::T lv[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const long e = (active && u < wb) ? off + (long)u * nl + t : zero;
        li[u] = (int)sidx[e];
        lv[u] = sval[e];
    }
    const int sr = active ? s0 + i : 0;
    const unsigned row = (unsigned)rowId[sr];
    const double rn = rowNorm[sr];
    const int myWave = (int)slotWave[sr];
    if (active && (t == 0 || iPrev != i)) rl0[i] = t;
    if (t == 0) rl0[nb] = nl;
    int rowL0 = 0, rowL1 = 0;
    wait();

    for (int k = k0; k < k1; ++k) {
        const size_t vo = (size_t)k * n;
        const XT g0 = ldX<FLOW>(x, vo, t < ne ? ex0 : 0u);
        const XT g1 = ldX<FLOW>(x, vo, t + BS < ne ? ex1 : 0u);
        const XT br = b[vo + (active ? row : 0u)];
        if (t < ne) xl[t] = g0;
        if (t + BS < ne) xl[t + BS] = g1;
        for (int c = t + 2 * BS; c < ne; c += BS) xl[c] = ldX<FLOW>(x, vo, (unsigned)ext[e0 + c]);
        __syncthreads();
        if (k == k0 && active) {
            rowL0 = rl0[i];
            rowL1 = rl0[i + 1];
        }
        for (int w = 0; w < nw; ++w) {
            // the lane's x entries, read once in one batch and kept
            // across the barrier (no other row of this wave writes them)
            XT cur[E];
            if (active && myWave == w) {
#pragma unroll
                for (int u = 0; u < E; ++u) cur[u] = xl[li[u]];
                XT dot = XT(0);
#pragma unroll
                for (int u = 0; u < E; ++u) dot += valOf<XT>(lv[u]) * cur[u];
                part[t] = dot;
            }
            __syncthreads();
            if (active && myWave == w) {
                XT rowDot = XT(0);
                for (int l = rowL0; l < rowL1; ++l) rowDot += part[l];
                const XT scale = (XT)(((double)rowDot - (double)br) / rn);
                // nonzero entries only: a padding entry may name a
                // column another row updates
#pragma unroll
                for (int u = 0; u < E; ++u) {
                    const XT a = valOf<XT>(lv[u]);
                    if (a != XT(0)) xl[li[u]] = cur[u] - scale * a;
                }
            }
            __syncthreads();
        }
        if (t < ne) stX<FLOW>(x, vo, ex0, xl[t]);
        if (t + BS < ne) stX<FLOW>(x, vo, ex1, xl[t + BS]);
        for (int c = t + 2 * BS; c < ne; c += BS) stX<FLOW>(x, vo, (unsigned)ext[e0 + c], xl[c]);
        __syncthreads();
    }
}

template <int BS, typename VT, typename XT, bool BATCH>
__global__ void __launch_bounds__(BS) k_kacz_bundle_vl(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const uint16_t* __restrict__ slotWave,
    const int* __restrict__ rowId, const double* __restrict__ rowNorm, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, long zero, XT* __restrict__ x0,
    const XT* __restrict__ b0, int n0, int nrhs, const int* __restrict__ chainPtr, int chain0, int nchain, int xcd,
    int ldsLocal, int ldsRows, const int* __restrict__ lanePtr, const uint16_t* __restrict__ laneSlot,
    XT* const* __restrict__ bX, const XT* const* __restrict__ bB, const int* __restrict__ bN) {
    extern __shared__ __align__(16) unsigned char smem[];
    XT* xl = reinterpret_cast<XT*>(smem);
    XT* part = xl + ldsLocal;                               // BS partial sums
    int* rl0 = reinterpret_cast<int*>(part + BS);           // row -> first lane
    (void)ldsRows;

    const int q = bundleOfBlock(nchain, xcd);
    if (q < 0) return;
    const int cLo = chainPtr[chain0 + q], cHi = chainPtr[chain0 + q + 1];
    const int k0 = (int)blockIdx.y;
    const int k1 = gridDim.y > 1 ? k0 + 1 : nrhs;
    auto runBundle = [&](const int bnd) {
        kaczBundleVL<BS, VT, XT, BATCH, false>(bnd, k0, k1, xl, part, rl0, NoWait{}, bSlot, bWave, bOff, bLen, bExt,
                                               slotWave, rowId, rowNorm, ext, sidx, sval, zero, x0, b0, n0, lanePtr,
                                               laneSlot, bX, bB, bN);
    };
    if (cHi == cLo + 1) {
        runBundle(cLo);
    } else {
        for (int bnd = cLo; bnd < cHi; ++bnd) {
            runBundle(bnd);
            __syncthreads();
        }
    }
}

// Kaczmarz with variable lanes per row, all phases of a sweep in one launch.
template <int BS, typename VT, typename XT, bool BATCH>
__global__ void __launch_bounds__(BS) k_kacz_flow_vl(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const uint16_t* __restrict__ slotWave,
    const int* __restrict__ rowId, const double* __restrict__ rowNorm, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, long zero, XT* x0, const XT* __restrict__ b0,
    int n0, int ldsLocal, const int* __restrict__ lanePtr, const uint16_t* __restrict__ laneSlot,
    XT* const* __restrict__ bX, const XT* const* __restrict__ bB, const int* __restrict__ bN,
    ReadyTables rt) {
    extern __shared__ __align__(16) unsigned char smem[];
    XT* xl = reinterpret_cast<XT*>(smem);
    XT* part = xl + ldsLocal;
    int* rl0 = reinterpret_cast<int*>(part + BS);
    const int k = (int)blockIdx.y;
    readyLoop<BS>(rt, [&](int bnd, auto wait) {
        kaczBundleVL<BS, VT, XT, BATCH, true>(bnd, k, k + 1, xl, part, rl0, wait, bSlot, bWave, bOff, bLen, bExt,
                                              slotWave, rowId, rowNorm, ext, sidx, sval, zero, x0, b0, n0, lanePtr,
                                              laneSlot, bX, bB, bN);
    });
}

// (lanes per row, register width) pairs with a uniform kernel
// instantiation; longer rows use variable lanes per row
int pickE(int tpr, int need) {
    static const int e1[] = {4, 8, 16}, e2[] = {8, 13, 16}, e4[] = {8, 16};
    if (tpr != 1 && tpr != 2 && tpr != 4) return -1;
    const int* c = tpr == 1 ? e1 : tpr == 2 ? e2 : e4;
    const int nc = tpr == 4 ? 2 : 3;
    for (int q = 0; q < nc; ++q)
        if (c[q] >= need) return c[q];
    return -1;
}

template <typename VT, typename XT>
struct BundleDev {
    BundleTable h;
    DevBuf<int> slot, wave, len, ext, extPtr, rowId, chainPtr, lanePtr;
    DevBuf<uint16_t> laneSlot, slotWave;
    DevBuf<long> off;
    DevBuf<uint16_t> sidx;
    DevBuf<VT> sval;
    long zero = 0;  // index of the zero entry in sidx / sval
    DevBuf<XT> diag;
    DevBuf<double> rowNorm;
    int load(BundleTable&& t) {
        h = std::move(t);
        int e;
        if ((e = upload(slot, h.slot.data(), h.slot.size()))) return e;
        if ((e = upload(wave, h.wave.data(), h.wave.size()))) return e;
        if ((e = upload(len, h.len.data(), h.len.size()))) return e;
        if ((e = upload(off, h.off.data(), h.off.size()))) return e;
        if ((e = upload(ext, h.ext.data(), h.ext.size()))) return e;
        if ((e = upload(extPtr, h.extPtr.data(), h.extPtr.size()))) return e;
        if ((e = upload(chainPtr, h.chainPtr.data(), h.chainPtr.size()))) return e;
        if (h.tpr == 0) {
            if ((e = upload(lanePtr, h.lanePtr.data(), h.lanePtr.size()))) return e;
            if ((e = upload(laneSlot, h.laneSlot.data(), h.laneSlot.size()))) return e;
        }
        if ((e = upload(rowId, h.rowId.data(), h.rowId.size()))) return e;
        // the zero entry: what a lane loads for an entry it does not have
        zero = (long)h.sidx.size();
        h.sidx.push_back(0);
        h.sval.push_back(0.0);
        if ((e = upload(sidx, h.sidx.data(), h.sidx.size()))) return e;
        std::vector<VT> v(h.sval.size());
        for (size_t q = 0; q < v.size(); ++q) v[q] = conv(h.sval[q]);
        if ((e = upload(sval, v.data(), v.size()))) return e;
        if (h.kind == RACE_K_KACZ) {
            if ((e = upload(rowNorm, h.rowNorm.data(), h.rowNorm.size()))) return e;
        } else {
            // Gauss-Seidel multiplies by the inverse diagonal
            std::vector<XT> dg(h.diag.size());
            for (size_t q = 0; q < dg.size(); ++q) dg[q] = (XT)(1.0 / h.diag[q]);
            if ((e = upload(diag, dg.data(), dg.size()))) return e;
        }
        // slot -> its wave inside the bundle (a bundle holds at most 8192 rows)
        std::vector<uint16_t> sw(h.rowId.size(), 0);
        for (size_t bd = 0; bd + 1 < h.wave.size(); ++bd)
            for (int w = h.wave[bd]; w < h.wave[bd + 1]; ++w)
                for (int sl = h.waveSlot[w]; sl < h.waveSlot[w + 1]; ++sl)
                    sw[sl] = (uint16_t)(w - h.wave[bd]);
        if ((e = upload(slotWave, sw.data(), sw.size()))) return e;
        // the host copies of the big arrays are no longer needed
        std::vector<uint16_t>().swap(h.sidx);
        std::vector<double>().swap(h.sval);
        return OK;
    }
    static VT conv(double v) {
        if constexpr (sizeof(VT) == 2)
            return toBf16(v);
        else
            return (VT)v;
    }
    // LDS bytes of phase p for workgroups of bs threads: the local column
    // space and, with variable lanes, a partial sum per lane and the first
    // lane of every row
    size_t lds(int p, int bs) const {
        const size_t var = h.tpr == 0 ? (size_t)bs * sizeof(XT) + (size_t)(h.phRows[p] + 1) * sizeof(int) : 0;
        return (size_t)h.phLocal[p] * sizeof(XT) + var + 32;
    }
};

// one fused or plain launch of the chains [c0, c0 + nchain) of phase p
template <int BS, int TPR, int EE, typename VT, typename XT, bool BATCH>
void launchBundles(int kind, int nrhs, const BundleDev<VT, XT>& d, int c0, int nchain, int p, int xcd, XT* x,
                   const XT* b, int n, XT* const* bX, const XT* const* bB, const int* bN, hipStream_t st) {
    const dim3 grid(xcd ? 8u * (unsigned)((nchain + 7) / 8) : (unsigned)nchain, (unsigned)std::max(1, nrhs));
    // chains of one bundle each: the kernel indexes bundles directly
    const std::vector<int>& cp = d.h.chainPtr;
    const int bundle0 = cp[c0 + nchain] - cp[c0] == nchain ? cp[c0] : -1;
    if (kind == RACE_K_KACZ)
        hipLaunchKernelGGL((k_kacz_bundle<BS, TPR, EE, VT, XT, BATCH>), grid, dim3(BS), d.lds(p, BS), st,
                           d.slot.p, d.wave.p, d.off.p, d.len.p, d.extPtr.p, d.slotWave.p, d.rowId.p,
                           d.rowNorm.p, d.ext.p, d.sidx.p, d.sval.p, d.zero, x, b, n, nrhs, d.chainPtr.p, c0, nchain, xcd,
                           d.h.phLocal[p], bX, bB, bN);
    else if (nrhs == 1 && (BS == 256 || BS == 512) && !BATCH && d.h.phExt[p] > 2 * BS)
        // many outside columns: gather up to 1536 (256 threads) or 2048
        // (512 threads) of them in one batch
        hipLaunchKernelGGL((k_gs_bundle<BS, TPR, EE, VT, XT, true, BATCH, (BS == 256 ? 6 : 4)>), dim3(grid.x), dim3(BS),
                           d.lds(p, BS),
                           st, d.slot.p, d.wave.p, d.off.p, d.len.p, d.extPtr.p, d.slotWave.p, d.rowId.p, d.diag.p,
                           d.ext.p, d.sidx.p, d.sval.p, d.zero, x, b, n, nrhs, d.chainPtr.p, c0, nchain, bundle0, xcd, bX,
                           bB, bN);
    else if (nrhs == 1)
        hipLaunchKernelGGL((k_gs_bundle<BS, TPR, EE, VT, XT, true, BATCH>), dim3(grid.x), dim3(BS), d.lds(p, BS), st,
                           d.slot.p, d.wave.p, d.off.p, d.len.p, d.extPtr.p, d.slotWave.p, d.rowId.p, d.diag.p,
                           d.ext.p, d.sidx.p, d.sval.p, d.zero, x, b, n, nrhs, d.chainPtr.p, c0, nchain, bundle0, xcd, bX,
                           bB, bN);
    else
        hipLaunchKernelGGL((k_gs_bundle<BS, TPR, EE, VT, XT, false, BATCH>), grid, dim3(BS), d.lds(p, BS),
                           st, d.slot.p, d.wave.p, d.off.p, d.len.p, d.extPtr.p, d.slotWave.p, d.rowId.p,
                           d.diag.p, d.ext.p, d.sidx.p, d.sval.p, d.zero, x, b, n, nrhs, d.chainPtr.p, c0, nchain, bundle0,
                           xcd, bX, bB, bN);
}

// the variable-lanes kernels (tpr == 0)
template <int BS, typename VT, typename XT, bool BATCH>
void launchBundlesVL(int kind, int nrhs, const BundleDev<VT, XT>& d, int c0, int nchain, int p, int xcd, XT* x,
                     const XT* b, int n, XT* const* bX, const XT* const* bB, const int* bN, hipStream_t st) {
    const dim3 grid(xcd ? 8u * (unsigned)((nchain + 7) / 8) : (unsigned)nchain, (unsigned)std::max(1, nrhs));
    if (kind == RACE_K_KACZ)
        hipLaunchKernelGGL((k_kacz_bundle_vl<BS, VT, XT, BATCH>), grid, dim3(BS), d.lds(p, BS), st, d.slot.p, d.wave.p,
                           d.off.p, d.len.p, d.extPtr.p, d.slotWave.p, d.rowId.p, d.rowNorm.p, d.ext.p, d.sidx.p,
                           d.sval.p, d.zero, x, b, n, nrhs, d.chainPtr.p, c0, nchain, xcd, d.h.phLocal[p], d.h.phRows[p],
                           d.lanePtr.p, d.laneSlot.p, bX, bB, bN);
    else
        hipLaunchKernelGGL((k_gs_bundle_vl<BS, VT, XT, BATCH>), grid, dim3(BS), d.lds(p, BS), st, d.slot.p, d.wave.p,
                           d.off.p, d.len.p, d.extPtr.p, d.slotWave.p, d.rowId.p, d.diag.p, d.ext.p, d.sidx.p,
                           d.sval.p, d.zero, x, b, n, nrhs, d.chainPtr.p, c0, nchain, xcd, d.h.phLocal[p], d.h.phRows[p],
                           d.lanePtr.p, d.laneSlot.p, bX, bB, bN);
}

template <typename VT, typename XT, bool BATCH>
void launchPhase(int bs, int tpr, int E, int kind, int nrhs, const BundleDev<VT, XT>& d, int c0, int nchain, int p,
                 int xcd, XT* x, const XT* b, int n, XT* const* bX, const XT* const* bB, const int* bN,
                 hipStream_t st) {
#define RACE_BUNDLE_CASE(BS_)                                                                              \
    switch (tpr * 100 + E) {                                                                               \
        case 104: launchBundles<BS_, 1, 4, VT, XT, BATCH>(kind, nrhs, d, c0, nchain, p, xcd, x, b, n, bX, bB, bN, st); break;   \
        case 108: launchBundles<BS_, 1, 8, VT, XT, BATCH>(kind, nrhs, d, c0, nchain, p, xcd, x, b, n, bX, bB, bN, st); break;   \
        case 116: launchBundles<BS_, 1, 16, VT, XT, BATCH>(kind, nrhs, d, c0, nchain, p, xcd, x, b, n, bX, bB, bN, st); break;  \
        case 208: launchBundles<BS_, 2, 8, VT, XT, BATCH>(kind, nrhs, d, c0, nchain, p, xcd, x, b, n, bX, bB, bN, st); break;   \
        case 213: launchBundles<BS_, 2, 13, VT, XT, BATCH>(kind, nrhs, d, c0, nchain, p, xcd, x, b, n, bX, bB, bN, st); break;  \
        case 216: launchBundles<BS_, 2, 16, VT, XT, BATCH>(kind, nrhs, d, c0, nchain, p, xcd, x, b, n, bX, bB, bN, st); break;  \
        case 416: launchBundles<BS_, 4, 16, VT, XT, BATCH>(kind, nrhs, d, c0, nchain, p, xcd, x, b, n, bX, bB, bN, st); break;  \
        case 16: launchBundlesVL<BS_, VT, XT, BATCH>(kind, nrhs, d, c0, nchain, p, xcd, x, b, n, bX, bB, bN, st); break;        \
        default: launchBundles<BS_, 4, 8, VT, XT, BATCH>(kind, nrhs, d, c0, nchain, p, xcd, x, b, n, bX, bB, bN, st); break;   \
    }
    switch (bs) {
        case 64: RACE_BUNDLE_CASE(64) break;
        case 128: RACE_BUNDLE_CASE(128) break;
        case 512: RACE_BUNDLE_CASE(512) break;
        case 1024: RACE_BUNDLE_CASE(1024) break;
        default: RACE_BUNDLE_CASE(256) break;
    }
#undef RACE_BUNDLE_CASE
}

// polls of one dependency flag or queue position before a launch gives up
// (about seconds; a correct dependency table never gets near it)
constexpr int kFlowPollLimit = 1 << 22;

// a bundle's prologue, gathers and stores, in dependency waves (weight of a
// bundle in the longest-path estimate of a readiness-driven sweep)
constexpr int kBundleFixedWaves = 4;

// Device state of the one-launch sweep of one direction: the bundle range
// and the dependencies of every chain, column ids with the coherent-read
// bit, and per right-hand side a flag per chain (tickets in chain order) or
// an arrival counter per chain and a ready queue (readiness-driven), and a
// ticket/epoch state.
struct FlowDev {
    DevBuf<int> depPtr, dep, flag, ext, lo, hi;
    DevBuf<FlowState> fs;
    // readiness-driven sweeps (Kaczmarz): successors, roots, arrival
    // counters and ready queues
    DevBuf<int> succPtr, succ, roots;
    DevBuf<unsigned> arrivals;
    DevBuf<unsigned long long> queue;
    bool readyMode = false;
    int nroots = 0;
    // readiness-driven: the longest dependency path of the sweep, in
    // dependency waves plus kBundleFixedWaves per bundle
    long longestPath = 0;
    int nchain = 0, grid = 0, nrhs = 1, maxLocal = 0, maxRows = 0;
    size_t lds = 0;
    // chains whose conflicts lie inside one phase (0 for a valid colouring)
    long samePhase = 0;
    // nkeys / colBase: the key space of the rows and columns (chainDeps);
    // kern: the kernel that will run it, for its occupancy
    // ready: readiness-driven (queue) instead of tickets in chain order
    template <typename VT, typename XT>
    int init(const BundleDev<VT, XT>& d, long nkeys, const std::vector<long>* colBase, int bs, int nrhs_,
             const void* kern, bool ready) {
        readyMode = ready;
        std::vector<int> dp, dd;
        samePhase = chainDeps(d.h, nkeys, colBase, &dp, &dd);
        nchain = (int)d.h.chainPtr.size() - 1;
        nrhs = std::max(1, nrhs_);
        for (int p = 0; p < d.h.numPhases(); ++p) {
            maxLocal = std::max(maxLocal, d.h.phLocal[p]);
            maxRows = std::max(maxRows, d.h.phRows[p]);
        }
        // one layout for every phase: the largest local column space, then
        // (variable lanes) the partial sums and the first lane of every row
        lds = (size_t)maxLocal * sizeof(XT) + 32 +
              (d.h.tpr == 0 ? (size_t)bs * sizeof(XT) + (size_t)(maxRows + 1) * sizeof(int) : 0);
        if (lds > kMaxLds) return OK;  // not used (the caller checks lds)
        // as many workgroups per right-hand side as are resident at once,
        // all sides together (more would only find the tickets gone)
        int per = 0, dev = 0, cus = 0;
        RACE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, bs, lds));
        RACE_HIP(hipGetDevice(&dev));
        RACE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int slots = std::max(1, std::max(per, 1) * cus / nrhs);
        grid = std::max(1, std::min(nchain, slots));
        // tickets in chain order (phase order); tables indexed by ticket
        // position
        std::vector<int> order(nchain);
        for (int c = 0; c < nchain; ++c) order[c] = c;
        std::vector<int> pos(nchain), hlo(nchain), hhi(nchain), qp(nchain + 1, 0), qd;
        for (int q = 0; q < nchain; ++q) pos[order[q]] = q;
        qd.reserve(dd.size());
        for (int q = 0; q < nchain; ++q) {
            const int c = order[q];
            hlo[q] = d.h.chainPtr[c];
            hhi[q] = d.h.chainPtr[c + 1];
            for (int j = dp[c]; j < dp[c + 1]; ++j) {
                if (pos[dd[j]] >= q) return ERR_INTERNAL;  // a dependency after its chain
                qd.push_back(pos[dd[j]]);
            }
            qp[q + 1] = (int)qd.size();
        }
        int e;
        if ((e = upload(depPtr, qp.data(), qp.size()))) return e;
        if ((e = upload(dep, qd.data(), qd.size()))) return e;
        if ((e = upload(lo, hlo.data(), hlo.size()))) return e;
        if ((e = upload(hi, hhi.data(), hhi.size()))) return e;
        const std::vector<int> fx = flowExt(d.h, nkeys, colBase);
        if ((e = upload(ext, fx.data(), fx.size()))) return e;
        const std::vector<int> zeros((size_t)std::max(nchain, 1) * nrhs, 0);
        if ((e = upload(flag, zeros.data(), zeros.size()))) return e;
        if (ready) {
            std::vector<int> sp(nchain + 1, 0), sv(qd.size()), rv;
            for (int c = 0; c < nchain; ++c)
                for (int j = qp[c]; j < qp[c + 1]; ++j) ++sp[qd[j] + 1];
            for (int c = 0; c < nchain; ++c) sp[c + 1] += sp[c];
            std::vector<int> fill(sp.begin(), sp.end() - 1);
            for (int c = 0; c < nchain; ++c)
                for (int j = qp[c]; j < qp[c + 1]; ++j) sv[fill[qd[j]]++] = c;
            for (int c = 0; c < nchain; ++c)
                if (qp[c + 1] == qp[c]) rv.push_back(c);
            nroots = (int)rv.size();
            // queue positions are chain numbers (order is the identity)
            for (long p : chainLongestPaths(d.h, dp, dd, kBundleFixedWaves)) longestPath = std::max(longestPath, p);
            if ((e = upload(succPtr, sp.data(), sp.size()))) return e;
            if ((e = upload(succ, sv.data(), sv.size()))) return e;
            if ((e = upload(roots, rv.data(), rv.size()))) return e;
            const std::vector<unsigned> az((size_t)std::max(nchain, 1) * nrhs, 0u);
            if ((e = upload(arrivals, az.data(), az.size()))) return e;
            // queue entries of epoch 0 never match a launch
            const std::vector<unsigned long long> qz((size_t)std::max(nchain, 1) * nrhs, 0ull);
            if ((e = upload(queue, qz.data(), qz.size()))) return e;
        }
        // flags start below the first epoch
        FlowState f0{};
        f0.epoch = 1;
        f0.tail = nroots;
        const std::vector<FlowState> fv(nrhs, f0);
        if ((e = upload(fs, fv.data(), fv.size()))) return e;
        return OK;
    }
    ReadyTables tables() const {
        ReadyTables r;
        r.lo = lo.p;
        r.hi = hi.p;
        r.depPtr = depPtr.p;
        r.succPtr = succPtr.p;
        r.succ = succ.p;
        r.roots = roots.p;
        r.arrivals = arrivals.p;
        r.queue = queue.p;
        r.fs = fs.p;
        r.nchain = nchain;
        r.nroots = nroots;
        r.pollLimit = kFlowPollLimit;
        return r;
    }
    int aborts(int* out) const {
        *out = 0;
        if (!fs.p) return OK;
        std::vector<FlowState> f(nrhs);
        RACE_HIP(hipMemcpy(f.data(), fs.p, sizeof(FlowState) * nrhs, hipMemcpyDeviceToHost));
        for (const FlowState& q : f) *out += q.aborts;
        return OK;
    }
};

template <int TPR, int EE, typename VT, typename XT>
int launchFlow(const BundleDev<VT, XT>& d, FlowDev& f, XT* x, const XT* b, hipStream_t st) {
    constexpr int BS = 256;
    auto kern = k_gs_flow<BS, TPR, EE, VT, XT>;
    hipLaunchKernelGGL(kern, dim3(f.grid), dim3(BS), f.lds, st, d.slot.p, d.wave.p, d.off.p, d.len.p, d.extPtr.p,
                       d.slotWave.p, d.rowId.p, d.diag.p, f.ext.p, d.sidx.p, d.sval.p, d.zero, x, b, f.lo.p, f.hi.p,
                       f.nchain, f.depPtr.p, f.dep.p, f.flag.p, f.fs.p, kFlowPollLimit);
    RACE_HIP(hipGetLastError());
    return OK;
}

template <typename VT, typename XT>
int launchFlowShape(int tpr, int E, const BundleDev<VT, XT>& d, FlowDev& f, XT* x, const XT* b, hipStream_t st) {
    switch (tpr * 100 + E) {
        case 104: return launchFlow<1, 4>(d, f, x, b, st);
        case 108: return launchFlow<1, 8>(d, f, x, b, st);
        case 116: return launchFlow<1, 16>(d, f, x, b, st);
        case 208: return launchFlow<2, 8>(d, f, x, b, st);
        case 213: return launchFlow<2, 13>(d, f, x, b, st);
        case 216: return launchFlow<2, 16>(d, f, x, b, st);
        case 408: return launchFlow<4, 8>(d, f, x, b, st);
        case 416: return launchFlow<4, 16>(d, f, x, b, st);
        default: return ERR_INTERNAL;
    }
}

// Kaczmarz one-launch sweep: grid (resident workgroups per side, nrhs)
template <int BS, int TPR, int EE, typename VT, typename XT, bool BATCH>
int launchKaczFlow(const BundleDev<VT, XT>& d, FlowDev& f, XT* x, const XT* b, int n, XT* const* bX,
                   const XT* const* bB, const int* bN, hipStream_t st) {
    if constexpr (TPR == 0) {
        hipLaunchKernelGGL((k_kacz_flow_vl<BS, VT, XT, BATCH>), dim3(f.grid, f.nrhs), dim3(BS), f.lds, st, d.slot.p,
                           d.wave.p, d.off.p, d.len.p, d.extPtr.p, d.slotWave.p, d.rowId.p, d.rowNorm.p, f.ext.p,
                           d.sidx.p, d.sval.p, d.zero, x, b, n, f.maxLocal, d.lanePtr.p, d.laneSlot.p, bX, bB, bN,
                           f.tables());
    } else {
        hipLaunchKernelGGL((k_kacz_flow<BS, TPR, EE, VT, XT, BATCH>), dim3(f.grid, f.nrhs), dim3(BS), f.lds, st,
                           d.slot.p, d.wave.p, d.off.p, d.len.p, d.extPtr.p, d.slotWave.p, d.rowId.p, d.rowNorm.p,
                           f.ext.p, d.sidx.p, d.sval.p, d.zero, x, b, n, bX, bB, bN, f.tables());
    }
    RACE_HIP(hipGetLastError());
    return OK;
}

template <typename VT, typename XT, bool BATCH>
int launchKaczFlowShape(int bs, int tpr, int E, const BundleDev<VT, XT>& d, FlowDev& f, XT* x, const XT* b, int n,
                        XT* const* bX, const XT* const* bB, const int* bN, hipStream_t st) {
#define RACE_KFLOW_CASE(BS_)                                                                                       \
    switch (tpr * 100 + E) {                                                                                       \
        case 104: return launchKaczFlow<BS_, 1, 4, VT, XT, BATCH>(d, f, x, b, n, bX, bB, bN, st);                  \
        case 108: return launchKaczFlow<BS_, 1, 8, VT, XT, BATCH>(d, f, x, b, n, bX, bB, bN, st);                  \
        case 116: return launchKaczFlow<BS_, 1, 16, VT, XT, BATCH>(d, f, x, b, n, bX, bB, bN, st);                 \
        case 208: return launchKaczFlow<BS_, 2, 8, VT, XT, BATCH>(d, f, x, b, n, bX, bB, bN, st);                  \
        case 213: return launchKaczFlow<BS_, 2, 13, VT, XT, BATCH>(d, f, x, b, n, bX, bB, bN, st);                 \
        case 216: return launchKaczFlow<BS_, 2, 16, VT, XT, BATCH>(d, f, x, b, n, bX, bB, bN, st);                 \
        case 408: return launchKaczFlow<BS_, 4, 8, VT, XT, BATCH>(d, f, x, b, n, bX, bB, bN, st);                  \
        case 416: return launchKaczFlow<BS_, 4, 16, VT, XT, BATCH>(d, f, x, b, n, bX, bB, bN, st);                 \
        case 16: return launchKaczFlow<BS_, 0, kVarE, VT, XT, BATCH>(d, f, x, b, n, bX, bB, bN, st);               \
        default: return ERR_INTERNAL;                                                                              \
    }
    switch (bs) {
        case 256: RACE_KFLOW_CASE(256)
        case 512: RACE_KFLOW_CASE(512)
        case 1024: RACE_KFLOW_CASE(1024)
        default: return ERR_INTERNAL;
    }
#undef RACE_KFLOW_CASE
}
// workgroup sizes the one-launch Kaczmarz sweep is built for
inline bool kaczFlowBs(int bs) { return bs == 256 || bs == 512 || bs == 1024; }

// the kernel a one-launch sweep of this shape runs (for its occupancy)
template <typename VT, typename XT>
const void* gsFlowKernel(int tpr, int E) {
    switch (tpr * 100 + E) {
        case 104: return reinterpret_cast<const void*>(k_gs_flow<256, 1, 4, VT, XT>);
        case 108: return reinterpret_cast<const void*>(k_gs_flow<256, 1, 8, VT, XT>);
        case 116: return reinterpret_cast<const void*>(k_gs_flow<256, 1, 16, VT, XT>);
        case 208: return reinterpret_cast<const void*>(k_gs_flow<256, 2, 8, VT, XT>);
        case 213: return reinterpret_cast<const void*>(k_gs_flow<256, 2, 13, VT, XT>);
        case 216: return reinterpret_cast<const void*>(k_gs_flow<256, 2, 16, VT, XT>);
        case 408: return reinterpret_cast<const void*>(k_gs_flow<256, 4, 8, VT, XT>);
        case 416: return reinterpret_cast<const void*>(k_gs_flow<256, 4, 16, VT, XT>);
        default: return nullptr;
    }
}
template <int BS, typename VT, typename XT, bool BATCH>
const void* kaczFlowKernelBs(int tpr, int E) {
    switch (tpr * 100 + E) {
        case 104: return reinterpret_cast<const void*>(k_kacz_flow<BS, 1, 4, VT, XT, BATCH>);
        case 108: return reinterpret_cast<const void*>(k_kacz_flow<BS, 1, 8, VT, XT, BATCH>);
        case 116: return reinterpret_cast<const void*>(k_kacz_flow<BS, 1, 16, VT, XT, BATCH>);
        case 208: return reinterpret_cast<const void*>(k_kacz_flow<BS, 2, 8, VT, XT, BATCH>);
        case 213: return reinterpret_cast<const void*>(k_kacz_flow<BS, 2, 13, VT, XT, BATCH>);
        case 216: return reinterpret_cast<const void*>(k_kacz_flow<BS, 2, 16, VT, XT, BATCH>);
        case 408: return reinterpret_cast<const void*>(k_kacz_flow<BS, 4, 8, VT, XT, BATCH>);
        case 416: return reinterpret_cast<const void*>(k_kacz_flow<BS, 4, 16, VT, XT, BATCH>);
        case 16: return reinterpret_cast<const void*>(k_kacz_flow_vl<BS, VT, XT, BATCH>);
        default: return nullptr;
    }
}
template <typename VT, typename XT, bool BATCH>
const void* kaczFlowKernel(int bs, int tpr, int E) {
    switch (bs) {
        case 256: return kaczFlowKernelBs<256, VT, XT, BATCH>(tpr, E);
        case 512: return kaczFlowKernelBs<512, VT, XT, BATCH>(tpr, E);
        case 1024: return kaczFlowKernelBs<1024, VT, XT, BATCH>(tpr, E);
        default: return nullptr;
    }
}

// lanes per row and register width for the longest row (diagonal excluded
// for Gauss-Seidel); E < 0 if no kernel instantiation fits.  Rows longer
// than 64 entries select variable lanes per row (tpr 0, kVarE registers
// per lane), so a few long rows do not widen every row of the matrix.
void pickShape(int kind, int maxRowLen, int forcedTpr, int* tpr, int* E) {
    int m = maxRowLen - (kind != RACE_K_KACZ ? 1 : 0);
    m = std::max(m, 1);
    if (forcedTpr <= 0 && m > 64) {
        *tpr = 0;
        *E = kVarE;
        return;
    }
    *tpr = forcedTpr > 0 ? forcedTpr : (m <= 16 ? 1 : m <= 32 ? 2 : 4);
    *E = pickE(*tpr, (m + *tpr - 1) / *tpr);
}

// rows per bundle: the workgroup's lanes divided by the lanes per row, or
// (variable lanes) the row limit alone, lanes checked by the builder
int rowsCapFor(int bundleRows, int bs, int tpr) { return tpr > 0 ? std::min(bundleRows, bs / tpr) : bundleRows; }

template <typename VT, typename XT>
class BundleFn : public DeviceFn {
  public:
    using Dev = BundleDev<VT, XT>;
    Dev df, dr;
    int nrhs = 1, n = 0, bs = 256, tpr = 2, E = 16, xcd = 1;
    std::unique_ptr<DeviceFn> fallback;
    // one launch per sweep (the "flow" option): Gauss-Seidel with one
    // right-hand side, 256-thread workgroups and uniform lanes per row;
    // Kaczmarz with 256-1024-thread workgroups
    bool flow = false;
    FlowDev ff, fr;

    int init(const race_kernel_desc& d, const ExecContext& ctx) {
        kind = d.kind;
        nrhs = d.nrhs > 0 ? d.nrhs : 1;
        n = d.nrow;
        x = d.x;
        b = d.b;
        bs = ctx.bundleThreads;
        xcd = ctx.xcdOrder;
        int maxLen = 0;
        for (int r = 0; r < n; ++r) maxLen = std::max(maxLen, d.rowPtr[r + 1] - d.rowPtr[r]);
        pickShape(kind, maxLen, ctx.bundleTpr, &tpr, &E);
        const int rowsCap = rowsCapFor(ctx.bundleRows, bs, tpr);
        const bool fp64 = std::is_same<VT, double>::value && std::is_same<XT, double>::value;
        auto fallBack = [&]() -> int {
            if (ctx.layout == 2 || !fp64) return ERR_INCOMPATIBILITY;  // bundles required
            return makeWaveFn(d, ctx, &fallback);
        };
        if (E < 0 || rowsCap < 1) return fallBack();
        BundleTable tf = buildBundles(*ctx.fwd, kind, n, d.rowPtr, d.col, d.val, rowsCap,
                                      ctx.bundleLocal, tpr, bs);
        if (!tf.ok) return fallBack();
        BundleTable tr = buildBundles(*ctx.rev, kind, n, d.rowPtr, d.col, d.val, rowsCap,
                                      ctx.bundleLocal, tpr, bs);
        if (!tr.ok) return fallBack();
        int e;
        if ((e = df.load(std::move(tf)))) return e;
        if ((e = dr.load(std::move(tr)))) return e;
        for (int p = 0; p < df.h.numPhases(); ++p)
            if (df.lds(p, bs) > kMaxLds) return fallBack();
        for (int p = 0; p < dr.h.numPhases(); ++p)
            if (dr.lds(p, bs) > kMaxLds) return fallBack();
        const bool gsFlow = kind == RACE_K_GS && nrhs == 1 && bs == 256 && tpr > 0;
        const bool kaczFlow = kind == RACE_K_KACZ && kaczFlowBs(bs);
        if (ctx.flow && (gsFlow || kaczFlow)) {
            const int sides = kind == RACE_K_KACZ ? nrhs : 1;
            const void* kern =
                kind == RACE_K_KACZ ? kaczFlowKernel<VT, XT, false>(bs, tpr, E) : gsFlowKernel<VT, XT>(tpr, E);
            if (!kern) return ERR_INTERNAL;
            const bool ready = kind == RACE_K_KACZ;
            if ((e = ff.init(df, n, nullptr, bs, sides, kern, ready)) ||
                (e = fr.init(dr, n, nullptr, bs, sides, kern, ready)))
                return e;
            // conflicts inside a phase would serialise its chains, and one
            // LDS layout serves every phase: keep the phase launches if
            // either does not hold
            flow = ff.samePhase == 0 && fr.samePhase == 0 && ff.lds <= kMaxLds && fr.lds <= kMaxLds;
        }
        return OK;
    }
    int flowAborts(int* out) const override {
        *out = 0;
        if (!flow) return OK;
        int a = 0, c = 0, e;
        if ((e = ff.aborts(&a)) || (e = fr.aborts(&c))) return e;
        *out = a + c;
        return OK;
    }

    int waves(bool rev) const override {
        if (fallback) return fallback->waves(rev);
        return (int)(rev ? dr.h.critical : df.h.critical);
    }
    bool graphable() const override { return fallback ? fallback->graphable() : true; }
    int launches(bool rev) const override {
        if (fallback) return fallback->launches(rev);
        if (flow) return 1;
        const std::vector<int>& pc = (rev ? dr : df).h.phaseChain;
        int k = 0;
        for (size_t p = 0; p + 1 < pc.size(); ++p) k += pc[p + 1] > pc[p];
        return k;
    }
    int run(bool rev, hipStream_t st) override {
        if (fallback) {
            fallback->x = x;
            fallback->b = b;
            return fallback->run(rev, st);
        }
        const Dev& d = rev ? dr : df;
        if (flow && kind == RACE_K_KACZ)
            return launchKaczFlowShape<VT, XT, false>(bs, tpr, E, d, rev ? fr : ff, (XT*)x, (const XT*)b, n, nullptr,
                                                      nullptr, nullptr, st);
        if (flow) return launchFlowShape<VT, XT>(tpr, E, d, rev ? fr : ff, (XT*)x, (const XT*)b, st);
        const std::vector<int>& pc = d.h.phaseChain;
        for (int p = 0; p + 1 < (int)pc.size(); ++p) {
            const int c0 = pc[p], nchain = pc[p + 1] - pc[p];
            if (nchain <= 0) continue;
            launchPhase<VT, XT, false>(bs, tpr, E, kind, nrhs, d, c0, nchain, p, xcd, (XT*)x, (const XT*)b, n,
                                       nullptr, nullptr, nullptr, st);
        }
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

// Gauss-Seidel or Kaczmarz on a batch of matrices, each with its own
// colouring and vectors: one fused launch per phase index runs the bundles
// of that phase of every matrix (concatBundles), so a batch of small
// matrices fills the device instead of issuing a few tiny launches per
// matrix and phase.  All matrices share the kernel shape (lanes per row,
// register width, workgroup size), chosen from the longest row.
template <typename VT, typename XT>
class BundleBatchFn : public DeviceFn {
  public:
    using Dev = BundleDev<VT, XT>;
    // matrices with the same kernel shape share fused tables
    struct Group {
        int tpr = 1, E = 4;
        Dev dev[2];
        DevBuf<XT*> bx[2];
        DevBuf<const XT*> bb[2];
        DevBuf<int> bn[2];
        FlowDev flow[2];
    };
    std::vector<std::unique_ptr<Group>> groups;
    int nrhs = 1, bs = 256, xcd = 0, maxPhases[2] = {0, 0};
    long critical[2] = {0, 0};
    // one launch per group and sweep (the "flow" option, Kaczmarz): every
    // matrix advances through its own phases, waiting only for its own
    // earlier chains, instead of every fused phase waiting for the slowest
    // matrix
    bool flow = false, flowOk = true;
    // flow: the groups hold different matrices, so their launches run
    // concurrently, the group with the longest dependency path first (on
    // the caller's stream), so that its long chains start at once and the
    // other groups (each on a stream of its own, forked from and joined back
    // into the caller's stream by events) fill the remaining workgroup slots
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> joinEv;
    hipEvent_t forkEv = nullptr;
    std::vector<int> launchOrder[2];

    ~BundleBatchFn() override {
        for (hipStream_t s : side) (void)hipStreamDestroy(s);
        for (hipEvent_t e : joinEv) (void)hipEventDestroy(e);
        if (forkEv) (void)hipEventDestroy(forkEv);
    }

    int init(const std::vector<BatchItem>& items, const ExecContext& ctx) {
        if (items.empty()) return ERR_INVALID_ARG;
        kind = items[0].d.kind;
        nrhs = items[0].d.nrhs > 0 ? items[0].d.nrhs : 1;
        bs = ctx.bundleThreads;
        xcd = ctx.xcdOrder;
        std::vector<int> shape(items.size());
        std::vector<std::pair<int, int>> shapes;
        for (size_t m = 0; m < items.size(); ++m) {
            const BatchItem& it = items[m];
            int maxLen = 0;
            for (int r = 0; r < it.d.nrow; ++r) maxLen = std::max(maxLen, it.d.rowPtr[r + 1] - it.d.rowPtr[r]);
            int tpr, E;
            pickShape(kind, maxLen, ctx.bundleTpr, &tpr, &E);
            if (E < 0 && ctx.bundleTpr > 0) pickShape(kind, maxLen, 0, &tpr, &E);  // row too long for the forced width
            if (E < 0 || rowsCapFor(ctx.bundleRows, bs, tpr) < 1) return ERR_INCOMPATIBILITY;
            const std::pair<int, int> key(tpr, E);
            auto f = std::find(shapes.begin(), shapes.end(), key);
            shape[m] = (int)(f - shapes.begin());
            if (f == shapes.end()) shapes.push_back(key);
        }
        for (size_t g = 0; g < shapes.size(); ++g) {
            auto grp = std::make_unique<Group>();
            grp->tpr = shapes[g].first;
            grp->E = shapes[g].second;
            const int rowsCap = rowsCapFor(ctx.bundleRows, bs, grp->tpr);
            for (int dir = 0; dir < 2; ++dir) {
                std::vector<int> members;
                for (size_t m = 0; m < items.size(); ++m)
                    if (shape[m] == (int)g) members.push_back((int)m);
                // the matrices' tables are independent: build them on host threads
                std::vector<BundleTable> tabs(members.size());
                std::atomic<size_t> next{0};
                std::atomic<bool> bad{false};
                auto work = [&]() {
                    for (size_t q; (q = next++) < members.size();) {
                        const BatchItem& it = items[members[q]];
                        tabs[q] = buildBundles(dir ? *it.rev : *it.fwd, kind, it.d.nrow, it.d.rowPtr, it.d.col,
                                               it.d.val, rowsCap, ctx.bundleLocal, grp->tpr, bs);
                        if (!tabs[q].ok) bad = true;
                    }
                };
                const size_t nth = std::min<size_t>(members.size(),
                                                    std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
                std::vector<std::thread> pool;
                for (size_t t = 1; t < nth; ++t) pool.emplace_back(work);
                work();
                for (std::thread& th : pool) th.join();
                if (bad) return ERR_INCOMPATIBILITY;
                std::vector<const BundleTable*> ptrs;
                for (size_t q = 0; q < members.size(); ++q) ptrs.push_back(&tabs[q]);
                std::vector<int> owner;
                BundleTable fused = concatBundles(ptrs, &owner);
                std::vector<BundleTable>().swap(tabs);
                std::vector<XT*> hx(owner.size());
                std::vector<const XT*> hb(owner.size());
                std::vector<int> hn(owner.size());
                for (size_t k = 0; k < owner.size(); ++k) {
                    const race_kernel_desc& d = items[members[owner[k]]].d;
                    hx[k] = (XT*)d.x;
                    hb[k] = (const XT*)d.b;
                    hn[k] = d.nrow;
                }
                int e;
                if ((e = upload(grp->bx[dir], hx.data(), hx.size()))) return e;
                if ((e = upload(grp->bb[dir], hb.data(), hb.size()))) return e;
                if ((e = upload(grp->bn[dir], hn.data(), hn.size()))) return e;
                if ((e = grp->dev[dir].load(std::move(fused)))) return e;
                for (int p = 0; p < grp->dev[dir].h.numPhases(); ++p)
                    if (grp->dev[dir].lds(p, bs) > kMaxLds) return ERR_INCOMPATIBILITY;
                maxPhases[dir] = std::max(maxPhases[dir], grp->dev[dir].h.numPhases());
                if (ctx.flow && kind == RACE_K_KACZ && kaczFlowBs(bs)) {
                    // every matrix its own range of row / column keys
                    std::vector<long> base(members.size() + 1, 0);
                    for (size_t q = 0; q < members.size(); ++q) base[q + 1] = base[q] + items[members[q]].d.nrow;
                    std::vector<long> colBase(owner.size());
                    for (size_t k = 0; k < owner.size(); ++k) colBase[k] = base[owner[k]];
                    const void* kern = kaczFlowKernel<VT, XT, true>(bs, grp->tpr, grp->E);
                    if (!kern) return ERR_INTERNAL;
                    if ((e = grp->flow[dir].init(grp->dev[dir], base.back(), &colBase, bs, nrhs, kern, true)))
                        return e;
                    // conflicts inside a phase, or one LDS layout for all
                    // phases too large: the fused phase launches then
                    if (grp->flow[dir].samePhase != 0 || grp->flow[dir].lds > kMaxLds) flowOk = false;
                }
            }
            groups.push_back(std::move(grp));
        }
        flow = flowOk && ctx.flow && kind == RACE_K_KACZ && kaczFlowBs(bs);
        for (int dir = 0; dir < 2; ++dir) {
            launchOrder[dir].resize(groups.size());
            for (size_t g = 0; g < groups.size(); ++g) launchOrder[dir][g] = (int)g;
            std::stable_sort(launchOrder[dir].begin(), launchOrder[dir].end(), [&](int a, int b) {
                return groups[a]->flow[dir].longestPath > groups[b]->flow[dir].longestPath;
            });
        }
        if (flow && groups.size() > 1) {
            RACE_HIP(hipEventCreateWithFlags(&forkEv, hipEventDisableTiming));
            for (size_t g = 1; g < groups.size(); ++g) {
                hipStream_t s = nullptr;
                hipEvent_t e = nullptr;
                RACE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
                side.push_back(s);
                RACE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                joinEv.push_back(e);
            }
        }
        for (int dir = 0; dir < 2; ++dir)
            for (int p = 0; p < maxPhases[dir]; ++p) {
                int w = 0;
                for (const auto& g : groups)
                    if (p < g->dev[dir].h.numPhases()) w = std::max(w, g->dev[dir].h.phWaves[p]);
                critical[dir] += w;
            }
        return OK;
    }
    int waves(bool rev) const override { return (int)critical[rev ? 1 : 0]; }
    int launches(bool rev) const override {
        if (flow) return (int)groups.size();
        int k = 0;
        const int dir = rev ? 1 : 0;
        for (int p = 0; p < maxPhases[dir]; ++p)
            for (const auto& g : groups)
                if (p < g->dev[dir].h.numPhases() && g->dev[dir].h.phaseChain[p + 1] > g->dev[dir].h.phaseChain[p]) ++k;
        return k;
    }
    int flowAborts(int* out) const override {
        *out = 0;
        if (!flow) return OK;
        for (const auto& g : groups)
            for (int dir = 0; dir < 2; ++dir) {
                int a = 0, e;
                if ((e = g->flow[dir].aborts(&a))) return e;
                *out += a;
            }
        return OK;
    }
    int run(bool rev, hipStream_t st) override {
        const int dir = rev ? 1 : 0;
        if (flow) {
            if (!side.empty()) RACE_HIP(hipEventRecord(forkEv, st));
            for (size_t q = 0; q < groups.size(); ++q) {
                Group& g = *groups[launchOrder[dir][q]];
                hipStream_t s = q == 0 ? st : side[q - 1];
                if (q > 0) RACE_HIP(hipStreamWaitEvent(s, forkEv, 0));
                int e = launchKaczFlowShape<VT, XT, true>(bs, g.tpr, g.E, g.dev[dir], g.flow[dir], nullptr, nullptr,
                                                          0, g.bx[dir].p, g.bb[dir].p, g.bn[dir].p, s);
                if (e) return e;
                if (q > 0) RACE_HIP(hipEventRecord(joinEv[q - 1], s));
            }
            for (hipEvent_t e : joinEv) RACE_HIP(hipStreamWaitEvent(st, e, 0));
            return OK;
        }
        // phase p of every group before phase p + 1 of any (stream order)
        for (int p = 0; p < maxPhases[dir]; ++p)
            for (const auto& g : groups) {
                const Dev& d = g->dev[dir];
                if (p >= d.h.numPhases()) continue;
                const int c0 = d.h.phaseChain[p], nchain = d.h.phaseChain[p + 1] - c0;
                if (nchain <= 0) continue;
                launchPhase<VT, XT, true>(bs, g->tpr, g->E, kind, nrhs, d, c0, nchain, p, xcd, nullptr, nullptr, 0,
                                          g->bx[dir].p, g->bb[dir].p, g->bn[dir].p, st);
            }
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

}  // namespace

int makeBundleBatchFn(const std::vector<BatchItem>& items, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out) {
    if (items.empty()) return ERR_INVALID_ARG;
    if (ctx.bundleThreads != 64 && ctx.bundleThreads != 128 && ctx.bundleThreads != 256 && ctx.bundleThreads != 512 &&
        ctx.bundleThreads != 1024)
        return ERR_INVALID_ARG;
    const race_kernel_desc& d0 = items[0].d;
    for (const BatchItem& it : items)
        if (it.d.kind != d0.kind || it.d.val_precision != d0.val_precision ||
            it.d.vec_precision != d0.vec_precision || it.d.nrhs != d0.nrhs || !it.fwd || !it.rev)
            return ERR_INVALID_ARG;
    if (d0.kind != RACE_K_GS && d0.kind != RACE_K_KACZ) return ERR_INVALID_ARG;
    int e;
    if (d0.vec_precision == RACE_FP64) {
        if (d0.val_precision != RACE_FP64) return ERR_NOT_IMPLEMENTED;
        auto f = std::make_unique<BundleBatchFn<double, double>>();
        e = f->init(items, ctx);
        *out = std::move(f);
    } else if (d0.val_precision == RACE_BF16) {
        auto f = std::make_unique<BundleBatchFn<uint16_t, float>>();
        e = f->init(items, ctx);
        *out = std::move(f);
    } else {
        auto f = std::make_unique<BundleBatchFn<float, float>>();
        e = f->init(items, ctx);
        *out = std::move(f);
    }
    return e;
}

int makeBundleFn(const race_kernel_desc& d, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out) {
    if (ctx.bundleThreads != 64 && ctx.bundleThreads != 128 && ctx.bundleThreads != 256 && ctx.bundleThreads != 512 &&
        ctx.bundleThreads != 1024)
        return ERR_INVALID_ARG;
    int e;
    if (d.vec_precision == RACE_FP64) {
        if (d.val_precision != RACE_FP64) return ERR_NOT_IMPLEMENTED;
        auto f = std::make_unique<BundleFn<double, double>>();
        e = f->init(d, ctx);
        *out = std::move(f);
    } else if (d.val_precision == RACE_BF16) {
        auto f = std::make_unique<BundleFn<uint16_t, float>>();
        e = f->init(d, ctx);
        *out = std::move(f);
    } else {
        auto f = std::make_unique<BundleFn<float, float>>();
        e = f->init(d, ctx);
        *out = std::move(f);
    }
    return e;
}

}  // namespace race
