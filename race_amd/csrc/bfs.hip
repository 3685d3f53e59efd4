// bfs.hip -- GPU frontier BFS with RACE's traversal semantics.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "bfs.hpp"

namespace race {

namespace {

constexpr int kBlock = 256;
constexpr int kHistBins = 4096;

__global__ void k_fill_i32(int* __restrict__ a, int v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) a[i] = v;
}

// seed a frontier with one row at level `lvl`
__global__ void k_seed(int* visit, int* dist, int* q, int row, int parent, int lvl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        visit[row] = parent;
        dist[row] = lvl;
        q[0] = row;
    }
}

// One BFS level: every frontier row claims its unvisited neighbours.
// counters[0] = next frontier length, counters[1] = in-range rows reached.
// bmask (POWER recursion stages, else null): rows of the boundary regions;
// such a row is claimed with the level and expanded like a range row but
// not counted (Traverse::TDStep src/traverse_GAP_graphSoA.cpp:184-197).
__global__ void __launch_bounds__(kBlock) k_bfs_step(
    const int* __restrict__ rowPtr, const int* __restrict__ col, const int* __restrict__ perm,
    const int* __restrict__ inv, int* __restrict__ visit, int* __restrict__ dist,
    const int* __restrict__ qin, int qlen, int* __restrict__ qout, int* __restrict__ counters,
    int lo, int hi, int lvl, int twoDist, int n, const unsigned char* __restrict__ bmask) {
    __shared__ int sCnt, sIn, sBase;
    if (threadIdx.x == 0) {
        sCnt = 0;
        sIn = 0;
    }
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long won = 0;  // claimed neighbours among the first 64 entries
    int cnt = 0, inCnt = 0, s = 0;
    if (t < qlen) {
        const int u = qin[t];
        const int pu = perm[u];
        s = rowPtr[pu];
        const int e = rowPtr[pu + 1];
        const bool uIn = (u >= lo && u < hi);
        for (int j = s; j < e; ++j) {
            int c = col[j];
            if (c >= n || c < 0) continue;  // remote column guard (child < graphSize)
            c = inv[c];
            if (visit[c] != -1) continue;
            const bool cIn = (c >= lo && c < hi);
            if (cIn) {
                if (atomicCAS(&visit[c], -1, u) == -1) {
                    dist[c] = lvl;
                    ++inCnt;
                    if (j - s < 64) {
                        won |= 1ull << (j - s);
                        ++cnt;
                    } else {
                        qout[atomicAdd(&counters[0], 1)] = c;
                    }
                }
            } else if (bmask && bmask[c]) {
                if (atomicCAS(&visit[c], -1, u) == -1) {
                    dist[c] = lvl;
                    if (j - s < 64) {
                        won |= 1ull << (j - s);
                        ++cnt;
                    } else {
                        qout[atomicAdd(&counters[0], 1)] = c;
                    }
                }
            } else if (twoDist && uIn) {
                if (atomicCAS(&visit[c], -1, u) == -1) {
                    if (j - s < 64) {
                        won |= 1ull << (j - s);
                        ++cnt;
                    } else {
                        qout[atomicAdd(&counters[0], 1)] = c;
                    }
                }
            }
        }
    }
    const int myOff = cnt ? atomicAdd(&sCnt, cnt) : 0;
    if (inCnt) atomicAdd(&sIn, inCnt);
    __syncthreads();
    if (threadIdx.x == 0) {
        sBase = sCnt ? atomicAdd(&counters[0], sCnt) : 0;
        if (sIn) atomicAdd(&counters[1], sIn);
    }
    __syncthreads();
    int w = sBase + myOff;
    while (won) {
        const int bit = __ffsll((long long)won) - 1;
        won &= won - 1;
        qout[w++] = inv[col[s + bit]];
    }
}

// ---------------------------------------------------------------------------
// Batched traversal: one workgroup per range, every level, island restart,
// level histogram and stable sort by level inside the workgroup, so a batch
// of disjoint ranges costs one launch (plus one for the composition) instead
// of one host round trip per level.  The result of a range depends only on
// the rows it holds and their order: neighbours outside it are tested for
// membership (distance 1) or reached by row identity (distance-2 bridges),
// so ranges of one batch are traversed concurrently and composed afterwards.
//
// Claims use a per-traversal stamp (visit[c] == stamp: reached) so no array
// is refilled between traversals; ranges of a batch are disjoint, so the
// shared visit array never sees two traversals on one row.  Distance-2
// bridge rows may be shared by several ranges: each workgroup marks them in
// its own n-entry slot of bvisit.
struct RangeJob {
    int lo, hi;
    long qOff, qCap;    // queue segment (range rows + bridge rows)
    long lvOff, lvCap;  // level arrays segment (at most 3 levels per row)
};

constexpr int kBatchMaxJobs = 4096;

// bitonic sort of BS 64-bit keys in LDS (blockDim.x == BS)
template <int BS>
__device__ void blockBitonic(unsigned long long* key) {
    const int t = threadIdx.x;
    for (int k = 2; k <= BS; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const int o = t ^ j;
            if (o > t) {
                const unsigned long long a = key[t], b = key[o];
                const bool up = (t & k) == 0;
                if ((a > b) == up) {
                    key[t] = b;
                    key[o] = a;
                }
            }
            __syncthreads();
        }
    }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_bfs_batch(
    const int* __restrict__ rowPtr, const int* __restrict__ col, const int* __restrict__ perm,
    const int* __restrict__ inv, int* __restrict__ visit, int* __restrict__ bvisit, int* __restrict__ dist,
    int* __restrict__ qAll, int* __restrict__ lvRow, int* __restrict__ lvNnz, int* __restrict__ lvStart,
    int* __restrict__ tmp, const RangeJob* __restrict__ jobs, int stamp0, int twoDist, int power, int n,
    int* __restrict__ nLev, int* __restrict__ err) {
    __shared__ int sNext, sIn, sRoot, sErr, sCarry;
    __shared__ unsigned long long sKey[BS];
    __shared__ int sScan[BS];
    const RangeJob jb = jobs[blockIdx.x];
    const int lo = jb.lo, hi = jb.hi, len = hi - lo;
    const int stamp = -2 - (stamp0 + (int)blockIdx.x);
    int* bv = twoDist ? bvisit + (size_t)blockIdx.x * n : nullptr;
    int* q = qAll + jb.qOff;
    const long qCap = jb.qCap;
    const int tid = threadIdx.x;

    // ---- traversal
    int head = 0, tail = 1, lvl = 0, prevIslandStop = lo, reached = 1;
    bool failed = false;
    if (tid == 0) {
        visit[lo] = stamp;
        dist[lo] = 0;
        q[0] = lo;
        sNext = 1;
        sIn = 0;
        sErr = 0;
        sRoot = 0x7fffffff;
    }
    __syncthreads();
    for (;;) {
        if (++lvl >= jb.lvCap) {
            failed = true;
            break;
        }
        for (int t = head + tid; t < tail; t += BS) {
            const int u = q[t];
            const int pu = perm[u];
            const int s = rowPtr[pu], e = rowPtr[pu + 1];
            const bool uIn = (u >= lo && u < hi);
            unsigned long long won = 0;
            int cnt = 0, inCnt = 0;
            for (int j = s; j < e; ++j) {
                int c = col[j];
                if (c >= n || c < 0) continue;
                c = inv[c];
                bool claimed = false;
                if (c >= lo && c < hi) {
                    if (visit[c] != stamp && atomicExch(&visit[c], stamp) != stamp) {
                        dist[c] = lvl;
                        ++inCnt;
                        claimed = true;
                    }
                } else if (twoDist && uIn) {
                    claimed = bv[c] != stamp && atomicExch(&bv[c], stamp) != stamp;
                }
                if (!claimed) continue;
                if (j - s < 64) {
                    won |= 1ull << (j - s);
                    ++cnt;
                } else {
                    const int w = atomicAdd(&sNext, 1);
                    if (w < qCap) q[w] = c;
                    else sErr = 1;
                }
            }
            if (inCnt) atomicAdd(&sIn, inCnt);
            int w = cnt ? atomicAdd(&sNext, cnt) : 0;
            while (won) {
                const int bit = __ffsll((long long)won) - 1;
                won &= won - 1;
                if (w < qCap) q[w] = inv[col[s + bit]];
                else sErr = 1;
                ++w;
            }
        }
        __syncthreads();
        head = tail;
        tail = sNext;
        reached += sIn;
        const int errNow = sErr;
        __syncthreads();  // every thread has read the counters
        if (tid == 0) sIn = 0;
        __syncthreads();
        if (errNow) {
            failed = true;
            break;
        }
        if (head < tail) continue;
        if (reached == len) break;
        // island: restart from the smallest row of the range not reached
        int root = 0x7fffffff;
        for (int c0 = prevIslandStop; c0 < hi; c0 += BS) {
            const int i = c0 + tid;
            // claims are device atomics: read them at L2, not from L1
            if (i < hi && __hip_atomic_load(&visit[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != stamp)
                atomicMin(&sRoot, i);
            __syncthreads();
            root = sRoot;
            __syncthreads();  // every thread has read sRoot before the next chunk
            if (root != 0x7fffffff) break;
        }
        if (root == 0x7fffffff || tail >= qCap) {
            failed = true;
            break;
        }
        if (!power) ++lvl;
        if (lvl >= jb.lvCap) {
            failed = true;
            break;
        }
        if (tid == 0) {
            visit[root] = stamp;
            dist[root] = lvl;
            q[tail] = root;
            sNext = tail + 1;
            sRoot = 0x7fffffff;
        }
        ++tail;
        ++reached;
        prevIslandStop = root;
        __syncthreads();
    }
    if (failed) {
        if (tid == 0) {
            nLev[blockIdx.x] = 0;
            atomicExch(err, 1);
        }
        return;
    }
    const int nL = lvl;
    int* hRow = lvRow + jb.lvOff;
    int* hNnz = lvNnz + jb.lvOff;
    int* start = lvStart + jb.lvOff;

    // ---- rows and graph entries per level
    for (int l = tid; l < nL; l += BS) {
        hRow[l] = 0;
        hNnz[l] = 0;
    }
    __syncthreads();
    for (int i = lo + tid; i < hi; i += BS) {
        const int d = dist[i];
        const int p = perm[i];
        atomicAdd(&hRow[d], 1);
        atomicAdd(&hNnz[d], rowPtr[p + 1] - rowPtr[p]);
    }
    __syncthreads();

    // ---- first position of every level (exclusive scan)
    if (tid == 0) sCarry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nL; c0 += BS) {
        const int v = (c0 + tid < nL) ? hRow[c0 + tid] : 0;
        sScan[tid] = v;
        __syncthreads();
        for (int o = 1; o < BS; o <<= 1) {
            const int add = tid >= o ? sScan[tid - o] : 0;
            __syncthreads();
            sScan[tid] += add;
            __syncthreads();
        }
        const int carry = sCarry;
        if (c0 + tid < nL) start[c0 + tid] = carry + sScan[tid] - v;
        __syncthreads();
        if (tid == BS - 1) sCarry = carry + sScan[BS - 1];
        __syncthreads();
    }

    // ---- stable placement by level, BS rows at a time in position order
    for (int c0 = lo; c0 < hi; c0 += BS) {
        const int i = c0 + tid;
        sKey[tid] = i < hi ? ((unsigned long long)(unsigned)dist[i] << 32) | (unsigned)tid : ~0ull;
        __syncthreads();
        blockBitonic<BS>(sKey);
        const unsigned long long k = sKey[tid];
        const bool valid = k != ~0ull;
        const unsigned d = (unsigned)(k >> 32);
        int r = 0;
        bool last = false;
        if (valid) {
            // first slot holding level d
            int a = 0, b = tid;
            const unsigned long long want = (unsigned long long)d << 32;
            while (a < b) {
                const int m = (a + b) >> 1;
                if (sKey[m] < want) a = m + 1;
                else b = m;
            }
            r = tid - a;
            const int src = (int)(k & 0xffffffffu);
            tmp[lo + start[d] + r] = perm[c0 + src];
            last = tid == BS - 1 || (unsigned)(sKey[tid + 1] >> 32) != d;
        }
        __syncthreads();  // every thread has read start[] and sKey
        if (last) start[d] += r + 1;
        __syncthreads();
    }
    if (tid == 0) nLev[blockIdx.x] = nL;
}

// perm/invPerm of every range of the batch from the sorted order, and the
// level arrays packed in job order (pack offset = levels of earlier jobs)
__global__ void __launch_bounds__(kBlock) k_compose_batch(
    const RangeJob* __restrict__ jobs, const int* __restrict__ nLev, const int* __restrict__ tmp,
    int* __restrict__ perm, int* __restrict__ inv, const int* __restrict__ lvRow,
    const int* __restrict__ lvNnz, int* __restrict__ packRow, int* __restrict__ packNnz) {
    __shared__ int sPart[kBlock];
    const RangeJob jb = jobs[blockIdx.x];
    for (int k = jb.lo + threadIdx.x; k < jb.hi; k += kBlock) {
        const int o = tmp[k];
        perm[k] = o;
        inv[o] = k;
    }
    int part = 0;
    for (int j = threadIdx.x; j < (int)blockIdx.x; j += kBlock) part += nLev[j];
    sPart[threadIdx.x] = part;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sPart[threadIdx.x] += sPart[threadIdx.x + o];
        __syncthreads();
    }
    const long base = sPart[0];
    const int nL = nLev[blockIdx.x];
    for (int l = threadIdx.x; l < nL; l += kBlock) {
        packRow[base + l] = lvRow[jb.lvOff + l];
        packNnz[base + l] = lvNnz[jb.lvOff + l];
    }
}

// smallest row in [start,hi) without a level
__global__ void k_fill_u8(unsigned char* __restrict__ a, unsigned char v, int lo, int hi) {
    int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (; i < hi; i += stride) a[i] = v;
}

// boundary rows no traversal reached get level `lvl` (createLevelData
// src/traverse_GAP_graphSoA.cpp:455-466); *flag = 1 if there was one
__global__ void k_untouched_to_level(int* __restrict__ dist, int lo, int hi, int lvl, int* __restrict__ flag) {
    int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    bool any = false;
    for (; i < hi; i += stride)
        if (dist[i] == -1) {
            dist[i] = lvl;
            any = true;
        }
    if (any) atomicOr(flag, 1);
}

__global__ void k_first_unvisited(const int* __restrict__ dist, int start, int hi, int* out) {
    int i = start + blockIdx.x * blockDim.x + threadIdx.x;
    int stride = gridDim.x * blockDim.x;
    int best = 0x7fffffff;
    for (; i < hi; i += stride)
        if (dist[i] == -1) {
            best = i;
            break;
        }
    if (best != 0x7fffffff) atomicMin(out, best);
}

// level histograms over [lo,hi): rows and graph entries per level
__global__ void __launch_bounds__(kBlock) k_level_hist(
    const int* __restrict__ dist, const int* __restrict__ perm, const int* __restrict__ rowPtr,
    int lo, int hi, int nLevels, int* __restrict__ levelRow, int* __restrict__ levelNnz) {
    __shared__ int hRow[kHistBins], hNnz[kHistBins];
    const bool local = nLevels <= kHistBins;
    if (local) {
        for (int i = threadIdx.x; i < nLevels; i += blockDim.x) {
            hRow[i] = 0;
            hNnz[i] = 0;
        }
        __syncthreads();
    }
    const long i0 = (long)lo + (long)blockIdx.x * blockDim.x * 8;
    for (int k = 0; k < 8; ++k) {
        const long i = i0 + (long)k * blockDim.x + threadIdx.x;
        if (i < hi) {
            const int d = dist[i];
            const int p = perm[i];
            const int len = rowPtr[p + 1] - rowPtr[p];
            if (local) {
                atomicAdd(&hRow[d], 1);
                atomicAdd(&hNnz[d], len);
            } else {
                atomicAdd(&levelRow[d], 1);
                atomicAdd(&levelNnz[d], len);
            }
        }
    }
    if (local) {
        __syncthreads();
        for (int l = threadIdx.x; l < nLevels; l += blockDim.x) {
            if (hRow[l]) {
                atomicAdd(&levelRow[l], hRow[l]);
                atomicAdd(&levelNnz[l], hNnz[l]);
            }
        }
    }
}

__global__ void k_sort_input(const int* __restrict__ dist, int lo, int len, int* __restrict__ keys,
                             int* __restrict__ vals) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < len) {
        keys[k] = dist[lo + k];
        vals[k] = lo + k;
    }
}

// perm_new[lo+k] = perm_old[order[k]]
__global__ void k_compose_gather(const int* __restrict__ perm, const int* __restrict__ order,
                                 int len, int* __restrict__ tmp) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < len) tmp[k] = perm[order[k]];
}

__global__ void k_compose_scatter(const int* __restrict__ tmp, int lo, int len,
                                  int* __restrict__ perm, int* __restrict__ inv) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < len) {
        const int o = tmp[k];
        perm[lo + k] = o;
        inv[o] = lo + k;
    }
}

inline unsigned gridFor(long n, int block = kBlock) {
    long g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

GpuTraverser::~GpuTraverser() {
    if (hPinned) (void)hipHostFree(hPinned);
}

int GpuTraverser::init(int n_, const int* rowPtr, const int* col, const int* initPerm,
                       const int* initInvPerm, hipStream_t stream) {
    n = n_;
    st = stream;
    nnz = rowPtr[n];
    hRowLen.resize(n);
    for (int i = 0; i < n; ++i) hRowLen[i] = rowPtr[i + 1] - rowPtr[i];
    hPerm.resize(n);
    hInv.resize(n);
    for (int i = 0; i < n; ++i) {
        hPerm[i] = initPerm ? initPerm[i] : i;
        hInv[i] = initInvPerm ? initInvPerm[i] : i;
    }
    RACE_HIP(dRowPtr.reserve(n + 1));
    RACE_HIP(dCol.reserve(nnz));
    RACE_HIP(dPerm.reserve(n));
    RACE_HIP(dInv.reserve(n));
    RACE_HIP(dVisit.reserve(n));
    RACE_HIP(dDist.reserve(n));
    RACE_HIP(dQ0.reserve(n));
    RACE_HIP(dQ1.reserve(n));
    RACE_HIP(dCount.reserve(4));
    RACE_HIP(dKeys.reserve(n));
    RACE_HIP(dKeysOut.reserve(n));
    RACE_HIP(dVals.reserve(n));
    RACE_HIP(dValsOut.reserve(n));
    RACE_HIP(dTmp.reserve(n));
    if (!hPinned) RACE_HIP(hipHostMalloc(reinterpret_cast<void**>(&hPinned), 4 * sizeof(int)));
    RACE_HIP(hipMemcpyAsync(dRowPtr.p, rowPtr, sizeof(int) * (n + 1), hipMemcpyHostToDevice, st));
    RACE_HIP(hipMemcpyAsync(dCol.p, col, sizeof(int) * nnz, hipMemcpyHostToDevice, st));
    RACE_HIP(hipMemcpyAsync(dPerm.p, hPerm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    RACE_HIP(hipMemcpyAsync(dInv.p, hInv.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    // no row carries a traversal stamp yet (stamps are <= -2)
    hipLaunchKernelGGL(k_fill_i32, dim3(gridFor(n / 4 + 1)), dim3(kBlock), 0, st, dVisit.p, -1, (long)n);
    RACE_HIP(hipGetLastError());
    RACE_HIP(hipStreamSynchronize(st));
    return OK;
}

int GpuTraverser::orderByRows(bool reverse, LevelData* out) {
    if (reverse) {
        std::reverse(hPerm.begin(), hPerm.end());
        for (int i = 0; i < n; ++i) hInv[hPerm[i]] = i;
        RACE_HIP(hipMemcpyAsync(dPerm.p, hPerm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        RACE_HIP(hipMemcpyAsync(dInv.p, hInv.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        RACE_HIP(hipStreamSynchronize(st));
    }
    out->levelRow.assign(n, 1);
    out->levelNnz.resize(n);
    for (int i = 0; i < n; ++i) out->levelNnz[i] = hRowLen[hPerm[i]];
    return OK;
}

int GpuTraverser::traverse(int dist, int lo, int hi, LevelData* out) {
    out->levelRow.clear();
    out->levelNnz.clear();
    if (lo >= hi) return OK;
    const int len = hi - lo;
    if (len <= rangeBlockMax) {
        std::vector<LevelData> one;
        int e = traverseBatch(dist, {{lo, hi}}, &one);
        if (e == OK) *out = std::move(one[0]);
        return e;
    }
    return traverseLevels(dist, lo, hi, out);
}

// A large range, level by level over the whole device.
int GpuTraverser::traverseLevels(int dist, int lo, int hi, LevelData* out) {
    auto t0 = std::chrono::steady_clock::now();
    int totalLevel = 0;
    int e = bfsLevels(dist, lo, hi, nullptr, &totalLevel);
    if (e) return e;
    if ((e = histogram(lo, hi, totalLevel, out))) return e;
    if ((e = sortByLevel(lo, hi, totalLevel))) return e;
    RACE_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    lastTraverseMs += std::chrono::duration<double, std::milli>(t1 - t0).count();
    ++traversals;
    return OK;
}

// The recursion stage's POWER traversal of [lo,hi) with boundary regions.
int GpuTraverser::traverseBoundary(int lo, int hi, const std::vector<std::pair<int, int>>& bounds,
                                   LevelData* out, std::vector<LevelData>* bout) {
    auto t0 = std::chrono::steady_clock::now();
    out->levelRow.clear();
    out->levelNnz.clear();
    bout->assign(bounds.size(), LevelData());
    if (lo >= hi) return OK;
    if (!dMask.p) {
        RACE_HIP(dMask.reserve(n));
        hipLaunchKernelGGL(k_fill_u8, dim3(gridFor(n / 4 + 1)), dim3(kBlock), 0, st, dMask.p, (unsigned char)0, 0, n);
    }
    for (const auto& b : bounds)
        if (b.second > b.first)
            hipLaunchKernelGGL(k_fill_u8, dim3(gridFor(b.second - b.first)), dim3(kBlock), 0, st, dMask.p,
                               (unsigned char)1, b.first, b.second);
    int totalLevel = 0;
    int e = bfsLevels(DIST_POWER, lo, hi, dMask.p, &totalLevel);
    for (const auto& b : bounds)
        if (b.second > b.first)
            hipLaunchKernelGGL(k_fill_u8, dim3(gridFor(b.second - b.first)), dim3(kBlock), 0, st, dMask.p,
                               (unsigned char)0, b.first, b.second);
    if (e) return e;
    // untouched boundary rows: one more level at the end; at least one level
    RACE_HIP(hipMemsetAsync(dCount.p + 3, 0, sizeof(int), st));
    for (const auto& b : bounds)
        if (b.second > b.first)
            hipLaunchKernelGGL(k_untouched_to_level, dim3(gridFor(b.second - b.first)), dim3(kBlock), 0, st,
                               dDist.p, b.first, b.second, totalLevel, dCount.p + 3);
    RACE_HIP(hipGetLastError());
    RACE_HIP(hipMemcpyAsync(hPinned + 3, dCount.p + 3, sizeof(int), hipMemcpyDeviceToHost, st));
    RACE_HIP(hipStreamSynchronize(st));
    if (hPinned[3]) totalLevel += 1;
    totalLevel = std::max(1, totalLevel);
    if ((e = histogram(lo, hi, totalLevel, out))) return e;
    for (size_t r = 0; r < bounds.size(); ++r)
        if ((e = histogram(bounds[r].first, bounds[r].second, totalLevel, &(*bout)[r]))) return e;
    // every region stably sorted by level, composed into the permutation
    if ((e = sortByLevel(lo, hi, totalLevel))) return e;
    for (const auto& b : bounds)
        if ((e = sortByLevel(b.first, b.second, totalLevel))) return e;
    RACE_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    lastTraverseMs += std::chrono::duration<double, std::milli>(t1 - t0).count();
    ++traversals;
    return OK;
}

// Level-synchronous BFS of [lo,hi) from lo with island restarts; dDist holds
// the levels afterwards (-1: not reached).
int GpuTraverser::bfsLevels(int dist, int lo, int hi, const unsigned char* bmask, int* levels) {
    const int len = hi - lo;
    // fresh visit/level state (Traverse ctor sets both to -1 for every row)
    hipLaunchKernelGGL(k_fill_i32, dim3(gridFor(n / 4 + 1)), dim3(kBlock), 0, st, dVisit.p, -1, (long)n);
    hipLaunchKernelGGL(k_fill_i32, dim3(gridFor(n / 4 + 1)), dim3(kBlock), 0, st, dDist.p, -1, (long)n);
    hipLaunchKernelGGL(k_seed, dim3(1), dim3(1), 0, st, dVisit.p, dDist.p, dQ0.p, lo, 0, 0);
    int* qin = dQ0.p;
    int* qout = dQ1.p;
    int qlen = 1;
    long reached = 1;
    int lvl = 0;
    int prevIslandStop = lo;
    const int two = (dist == DIST_TWO) ? 1 : 0;
    while (qlen > 0) {
        lvl += 1;
        RACE_HIP(hipMemsetAsync(dCount.p, 0, 2 * sizeof(int), st));
        hipLaunchKernelGGL(k_bfs_step, dim3(gridFor(qlen)), dim3(kBlock), 0, st, dRowPtr.p, dCol.p,
                           dPerm.p, dInv.p, dVisit.p, dDist.p, qin, qlen, qout, dCount.p, lo, hi,
                           lvl, two, n, bmask);
        RACE_HIP(hipGetLastError());
        RACE_HIP(hipMemcpyAsync(hPinned, dCount.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        RACE_HIP(hipStreamSynchronize(st));
        qlen = hPinned[0];
        reached += hPinned[1];
        std::swap(qin, qout);
        if (qlen == 0 && reached != len) {
            // island: restart from the smallest row without a level
            hPinned[2] = 0x7fffffff;
            RACE_HIP(hipMemcpyAsync(dCount.p + 2, hPinned + 2, sizeof(int), hipMemcpyHostToDevice, st));
            int span = hi - prevIslandStop;
            hipLaunchKernelGGL(k_first_unvisited, dim3(gridFor(span / 16 + 1)), dim3(kBlock), 0, st,
                               dDist.p, prevIslandStop, hi, dCount.p + 2);
            RACE_HIP(hipMemcpyAsync(hPinned + 2, dCount.p + 2, sizeof(int), hipMemcpyDeviceToHost, st));
            RACE_HIP(hipStreamSynchronize(st));
            const int root = hPinned[2];
            if (root == 0x7fffffff) {
                std::fprintf(stderr, "race_amd: traversal lost rows in [%d,%d)\n", lo, hi);
                return ERR_GRAPH_TRAVERSAL;
            }
            if (dist != DIST_POWER) lvl += 1;
            hipLaunchKernelGGL(k_seed, dim3(1), dim3(1), 0, st, dVisit.p, dDist.p, qin, root, root, lvl);
            qlen = 1;
            reached += 1;
            prevIslandStop = root;
        }
    }
    *levels = lvl;
    return OK;
}

// levelRow / levelNnz of the rows [lo,hi) over totalLevel levels
// (findLevelData src/traverse_GAP_graphSoA.cpp:385-447)
int GpuTraverser::histogram(int lo, int hi, int totalLevel, LevelData* out) {
    const int len = hi - lo;
    out->levelRow.assign(totalLevel, 0);
    out->levelNnz.assign(totalLevel, 0);
    if (len <= 0) return OK;
    RACE_HIP(dLevelRow.reserve(totalLevel + 1));
    RACE_HIP(dLevelNnz.reserve(totalLevel + 1));
    RACE_HIP(hipMemsetAsync(dLevelRow.p, 0, sizeof(int) * totalLevel, st));
    RACE_HIP(hipMemsetAsync(dLevelNnz.p, 0, sizeof(int) * totalLevel, st));
    hipLaunchKernelGGL(k_level_hist, dim3(gridFor(len, kBlock * 8)), dim3(kBlock), 0, st, dDist.p,
                       dPerm.p, dRowPtr.p, lo, hi, totalLevel, dLevelRow.p, dLevelNnz.p);
    RACE_HIP(hipGetLastError());
    out->levelRow.resize(totalLevel);
    out->levelNnz.resize(totalLevel);
    RACE_HIP(hipMemcpyAsync(out->levelRow.data(), dLevelRow.p, sizeof(int) * totalLevel,
                            hipMemcpyDeviceToHost, st));
    RACE_HIP(hipMemcpyAsync(out->levelNnz.data(), dLevelNnz.p, sizeof(int) * totalLevel,
                            hipMemcpyDeviceToHost, st));
    RACE_HIP(hipStreamSynchronize(st));
    return OK;
}

// stable sort of [lo,hi) by level (LSD radix sort is stable), composed into
// perm / invPerm on the device and the host mirror
int GpuTraverser::sortByLevel(int lo, int hi, int totalLevel) {
    const int len = hi - lo;
    if (len <= 0) return OK;
    hipLaunchKernelGGL(k_sort_input, dim3(gridFor(len)), dim3(kBlock), 0, st, dDist.p, lo, len,
                       dKeys.p, dVals.p);
    int endBit = 1;
    while ((1l << endBit) <= (long)totalLevel) ++endBit;
    size_t tmpBytes = 0;
    RACE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpBytes, dKeys.p, dKeysOut.p, dVals.p,
                                                dValsOut.p, len, 0, endBit, st));
    RACE_HIP(dSortTmp.reserve(tmpBytes));
    RACE_HIP(hipcub::DeviceRadixSort::SortPairs(dSortTmp.p, tmpBytes, dKeys.p, dKeysOut.p, dVals.p,
                                                dValsOut.p, len, 0, endBit, st));
    hipLaunchKernelGGL(k_compose_gather, dim3(gridFor(len)), dim3(kBlock), 0, st, dPerm.p,
                       dValsOut.p, len, dTmp.p);
    hipLaunchKernelGGL(k_compose_scatter, dim3(gridFor(len)), dim3(kBlock), 0, st, dTmp.p, lo, len,
                       dPerm.p, dInv.p);
    RACE_HIP(hipGetLastError());
    // host mirror of the range
    RACE_HIP(hipMemcpyAsync(hPerm.data() + lo, dPerm.p + lo, sizeof(int) * len,
                            hipMemcpyDeviceToHost, st));
    RACE_HIP(hipStreamSynchronize(st));
    for (int i = lo; i < hi; ++i) hInv[hPerm[i]] = i;
    return OK;
}

int GpuTraverser::traverseBatch(int dist, const std::vector<std::pair<int, int>>& ranges,
                                std::vector<LevelData>* outs) {
    outs->assign(ranges.size(), LevelData());
    std::vector<size_t> small;
    for (size_t r = 0; r < ranges.size(); ++r) {
        const int len = ranges[r].second - ranges[r].first;
        if (len <= 0) continue;
        if (len > rangeBlockMax) {
            int e = traverseLevels(dist, ranges[r].first, ranges[r].second, &(*outs)[r]);
            if (e != OK) return e;
        } else {
            small.push_back(r);
        }
    }
    if (small.empty()) return OK;
    auto t0 = std::chrono::steady_clock::now();
    const int two = (dist == DIST_TWO) ? 1 : 0;
    // distance-2 bridge slots: one n-entry array per concurrent range, at
    // most 1 GiB of them
    int maxJobs = kBatchMaxJobs;
    if (two) {
        const long slots = std::max(1l, std::min((long)kBatchMaxJobs, (1l << 28) / std::max(1, n)));
        maxJobs = (int)slots;
        if ((long)bvisitSlots < slots) {
            RACE_HIP(dBVisit.reserve((size_t)slots * n));
            hipLaunchKernelGGL(k_fill_i32, dim3(gridFor(slots * n / 8 + 1)), dim3(kBlock), 0, st, dBVisit.p,
                               -1, slots * (long)n);
            RACE_HIP(hipGetLastError());
            bvisitSlots = (int)slots;
        }
    }
    for (size_t c0 = 0; c0 < small.size(); c0 += maxJobs) {
        const size_t nj = std::min(small.size() - c0, (size_t)maxJobs);
        std::vector<RangeJob> jobs(nj);
        long qTotal = 0, lvTotal = 0;
        int spanLo = n, spanHi = 0, maxLen = 0;
        for (size_t j = 0; j < nj; ++j) {
            const int lo = ranges[small[c0 + j]].first, hi = ranges[small[c0 + j]].second;
            const int len = hi - lo;
            long bridges = 0;
            if (two) {
                for (int i = lo; i < hi; ++i) bridges += hRowLen[hPerm[i]];
                bridges = std::min(bridges, (long)(n - len));
            }
            jobs[j] = RangeJob{lo, hi, qTotal, len + bridges, lvTotal, 3l * len + 2};
            qTotal += len + bridges;
            lvTotal += 3l * len + 2;
            spanLo = std::min(spanLo, lo);
            spanHi = std::max(spanHi, hi);
            maxLen = std::max(maxLen, len);
        }
        RACE_HIP(dJobs.reserve(nj * sizeof(RangeJob)));
        RACE_HIP(dBatchQ.reserve(qTotal));
        RACE_HIP(dLvRow.reserve(lvTotal));
        RACE_HIP(dLvNnz.reserve(lvTotal));
        RACE_HIP(dLvStart.reserve(lvTotal));
        RACE_HIP(dPackRow.reserve(lvTotal));
        RACE_HIP(dPackNnz.reserve(lvTotal));
        RACE_HIP(dNLev.reserve(nj + 1));
        RACE_HIP(hipMemcpyAsync(dJobs.p, jobs.data(), nj * sizeof(RangeJob), hipMemcpyHostToDevice, st));
        RACE_HIP(hipMemsetAsync(dNLev.p + nj, 0, sizeof(int), st));
        const RangeJob* dj = reinterpret_cast<const RangeJob*>(dJobs.p);
        const int stamp0 = stampNext;
        stampNext += (int)nj;
        if (maxLen <= 4096)
            hipLaunchKernelGGL(k_bfs_batch<256>, dim3(nj), dim3(256), 0, st, dRowPtr.p, dCol.p, dPerm.p,
                               dInv.p, dVisit.p, dBVisit.p, dDist.p, dBatchQ.p, dLvRow.p, dLvNnz.p, dLvStart.p,
                               dTmp.p, dj, stamp0, two, dist == DIST_POWER ? 1 : 0, n, dNLev.p, dNLev.p + nj);
        else
            hipLaunchKernelGGL(k_bfs_batch<1024>, dim3(nj), dim3(1024), 0, st, dRowPtr.p, dCol.p, dPerm.p,
                               dInv.p, dVisit.p, dBVisit.p, dDist.p, dBatchQ.p, dLvRow.p, dLvNnz.p, dLvStart.p,
                               dTmp.p, dj, stamp0, two, dist == DIST_POWER ? 1 : 0, n, dNLev.p, dNLev.p + nj);
        RACE_HIP(hipGetLastError());
        hNLev.resize(nj + 1);
        RACE_HIP(hipMemcpyAsync(hNLev.data(), dNLev.p, (nj + 1) * sizeof(int), hipMemcpyDeviceToHost, st));
        RACE_HIP(hipStreamSynchronize(st));
        if (hNLev[nj]) {
            std::fprintf(stderr, "race_amd: batched traversal of %zu ranges lost rows\n", nj);
            return ERR_GRAPH_TRAVERSAL;
        }
        hipLaunchKernelGGL(k_compose_batch, dim3(nj), dim3(kBlock), 0, st, dj, dNLev.p, dTmp.p, dPerm.p, dInv.p,
                           dLvRow.p, dLvNnz.p, dPackRow.p, dPackNnz.p);
        RACE_HIP(hipGetLastError());
        long packed = 0;
        for (size_t j = 0; j < nj; ++j) packed += hNLev[j];
        hPackRow.resize(packed);
        hPackNnz.resize(packed);
        RACE_HIP(hipMemcpyAsync(hPackRow.data(), dPackRow.p, packed * sizeof(int), hipMemcpyDeviceToHost, st));
        RACE_HIP(hipMemcpyAsync(hPackNnz.data(), dPackNnz.p, packed * sizeof(int), hipMemcpyDeviceToHost, st));
        RACE_HIP(hipMemcpyAsync(hPerm.data() + spanLo, dPerm.p + spanLo, sizeof(int) * (spanHi - spanLo),
                                hipMemcpyDeviceToHost, st));
        RACE_HIP(hipStreamSynchronize(st));
        long off = 0;
        for (size_t j = 0; j < nj; ++j) {
            LevelData& ld = (*outs)[small[c0 + j]];
            ld.levelRow.assign(hPackRow.begin() + off, hPackRow.begin() + off + hNLev[j]);
            ld.levelNnz.assign(hPackNnz.begin() + off, hPackNnz.begin() + off + hNLev[j]);
            off += hNLev[j];
            for (int i = jobs[j].lo; i < jobs[j].hi; ++i) hInv[hPerm[i]] = i;
        }
        traversals += (long)nj;
    }
    auto t1 = std::chrono::steady_clock::now();
    lastTraverseMs += std::chrono::duration<double, std::milli>(t1 - t0).count();
    return OK;
}

}  // namespace race
