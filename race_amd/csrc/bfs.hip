// bfs.hip -- GPU frontier BFS with RACE's traversal semantics.
#include <hipcub/hipcub.hpp>

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
__global__ void __launch_bounds__(kBlock) k_bfs_step(
    const int* __restrict__ rowPtr, const int* __restrict__ col, const int* __restrict__ perm,
    const int* __restrict__ inv, int* __restrict__ visit, int* __restrict__ dist,
    const int* __restrict__ qin, int qlen, int* __restrict__ qout, int* __restrict__ counters,
    int lo, int hi, int lvl, int twoDist, int n) {
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

// smallest row in [start,hi) without a level
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
    RACE_HIP(hipStreamSynchronize(st));
    return OK;
}

int GpuTraverser::traverse(int dist, int lo, int hi, LevelData* out) {
    auto t0 = std::chrono::steady_clock::now();
    out->levelRow.clear();
    out->levelNnz.clear();
    if (lo >= hi) return OK;
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
                           lvl, two, n);
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
    const int totalLevel = lvl;

    // histograms
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

    // stable sort of the range by level (LSD radix sort is stable)
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
    auto t1 = std::chrono::steady_clock::now();
    lastTraverseMs += std::chrono::duration<double, std::milli>(t1 - t0).count();
    ++traversals;
    return OK;
}

}  // namespace race
