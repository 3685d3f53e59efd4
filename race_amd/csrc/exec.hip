// exec.hip -- colour-phase launcher: SymmSpMV, Gauss-Seidel, Kaczmarz, SpMV
// and the level-blocked matrix power kernel on the MI355X.
#include <algorithm>
#include <cstring>

#include "bundle.hpp"
#include "exec.hpp"

namespace race {

// ---------------------------------------------------------------------------
// dependency waves inside leaves
// ---------------------------------------------------------------------------
WaveTable buildWaves(const Schedule& s, int kind, int n, const int* rp, const int* col) {
    WaveTable t;
    t.phaseLeafPtr = s.phasePtr;
    t.leafWavePtr.push_back(0);
    t.wavePtr.push_back(0);
    t.rows.reserve(n);
    std::vector<int> depth(n, 0), W(n, 0), R(n, 0), stamp(n, -1);
    std::vector<int> bucket;
    const int nLeaves = (int)s.leafLo.size();
    for (int p = 0; p + 1 < (int)s.phasePtr.size(); ++p) {
        int phaseMax = 0;
        for (int l = s.phasePtr[p]; l < s.phasePtr[p + 1]; ++l) {
            const int lo = s.leafLo[l], hi = s.leafHi[l];
            int maxD = 0;
            for (int k = 0; k < hi - lo; ++k) {
                const int r = s.rev ? hi - 1 - k : lo + k;
                int d = 0;
                if (kind == RACE_K_KACZ) {
                    for (int j = rp[r]; j < rp[r + 1]; ++j) {
                        const int c = col[j];
                        if (stamp[c] == l) d = std::max(d, W[c]);
                    }
                    ++d;
                    for (int j = rp[r]; j < rp[r + 1]; ++j) {
                        const int c = col[j];
                        stamp[c] = l;
                        W[c] = d;
                    }
                } else {
                    // true dependency on earlier writers of what r reads,
                    // anti dependency on earlier readers of what r writes
                    for (int j = rp[r]; j < rp[r + 1]; ++j) {
                        const int c = col[j];
                        if (c != r && stamp[c] == l) d = std::max(d, W[c]);
                    }
                    if (stamp[r] == l) d = std::max(d, R[r]);
                    ++d;
                    if (stamp[r] != l) {
                        stamp[r] = l;
                        R[r] = 0;
                    }
                    W[r] = d;
                    for (int j = rp[r]; j < rp[r + 1]; ++j) {
                        const int c = col[j];
                        if (c == r) continue;
                        if (stamp[c] != l) {
                            stamp[c] = l;
                            W[c] = 0;
                            R[c] = 0;
                        }
                        R[c] = std::max(R[c], d);
                    }
                }
                depth[r] = d;
                maxD = std::max(maxD, d);
            }
            // rows grouped by depth, ascending row id inside a wave
            bucket.assign(maxD + 2, 0);
            for (int r = lo; r < hi; ++r) bucket[depth[r] + 1]++;
            for (int w = 1; w <= maxD + 1; ++w) bucket[w] += bucket[w - 1];
            const size_t base = t.rows.size();
            t.rows.resize(base + (hi - lo));
            for (int r = lo; r < hi; ++r) t.rows[base + bucket[depth[r]]++] = r;
            // bucket[w] now holds the end of wave w (1-based depths)
            for (int w = 1; w <= maxD; ++w) t.wavePtr.push_back((int)base + bucket[w]);
            t.leafWavePtr.push_back((int)t.wavePtr.size() - 1);
            phaseMax = std::max(phaseMax, maxD);
        }
        t.critical += phaseMax;
    }
    (void)nLeaves;
    return t;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
namespace {

constexpr int kBlock = 256;
constexpr size_t kMaxLds = 64 * 1024;  // LDS bytes one workgroup may allocate
constexpr int kBatch = 16;             // SELL entries loaded per batch in bundle prologues

template <int W>
__device__ __forceinline__ double groupSum(double v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// Gauss-Seidel: one workgroup per leaf, waves in order, TPR lanes per row.
// x[row] = (b[row] - sum_{j != diag} a_j x[c_j]) / a_diag, diagonal first.
template <int TPR>
__global__ void __launch_bounds__(kBlock) k_gs_leaf(
    const int* __restrict__ rp, const int* __restrict__ col, const double* __restrict__ val,
    double* x, const double* __restrict__ b, int n, int nrhs, const int* __restrict__ leafWave,
    const int* __restrict__ wavePtr, const int* __restrict__ rows, int leaf0) {
    const int leaf = leaf0 + blockIdx.x;
    const int w0 = leafWave[leaf], w1 = leafWave[leaf + 1];
    const int lane = threadIdx.x % TPR, grp = threadIdx.x / TPR, ngrp = blockDim.x / TPR;
    for (int w = w0; w < w1; ++w) {
        const int r0 = wavePtr[w], r1 = wavePtr[w + 1];
        for (int i = r0 + grp; i < r1; i += ngrp) {
            const int row = rows[i];
            const int s = rp[row], e = rp[row + 1];
            for (int k = 0; k < nrhs; ++k) {
                const size_t off = (size_t)k * n;
                double t = 0.0;
                for (int j = s + 1 + lane; j < e; j += TPR) t += val[j] * x[off + col[j]];
                t = groupSum<TPR>(t);
                if (lane == 0) x[off + row] = (b[off + row] - t) / val[s];
            }
        }
        __syncthreads();
    }
}

// Kaczmarz: x += (b_r - a_r.x)/|a_r|^2 a_r, rows of a wave share no column.
template <int TPR>
__global__ void __launch_bounds__(kBlock) k_kacz_leaf(
    const int* __restrict__ rp, const int* __restrict__ col, const double* __restrict__ val,
    double* x, const double* __restrict__ b, int n, int nrhs, const int* __restrict__ leafWave,
    const int* __restrict__ wavePtr, const int* __restrict__ rows, int leaf0) {
    const int leaf = leaf0 + blockIdx.x;
    const int w0 = leafWave[leaf], w1 = leafWave[leaf + 1];
    const int lane = threadIdx.x % TPR, grp = threadIdx.x / TPR, ngrp = blockDim.x / TPR;
    for (int w = w0; w < w1; ++w) {
        const int r0 = wavePtr[w], r1 = wavePtr[w + 1];
        for (int i = r0 + grp; i < r1; i += ngrp) {
            const int row = rows[i];
            const int s = rp[row], e = rp[row + 1];
            for (int k = 0; k < nrhs; ++k) {
                const size_t off = (size_t)k * n;
                double dot = 0.0, nrm = 0.0;
                for (int j = s + lane; j < e; j += TPR) {
                    const double m = val[j];
                    dot += m * x[off + col[j]];
                    nrm += m * m;
                }
                dot = groupSum<TPR>(dot);
                nrm = groupSum<TPR>(nrm);
                const double scale = (dot - b[off + row]) / nrm;
                for (int j = s + lane; j < e; j += TPR) x[off + col[j]] -= scale * val[j];
            }
        }
        __syncthreads();
    }
}

// SymmSpMV over row chunks of one phase: b += (U + U^T - D) x with the
// transpose part scattered by FP64 atomics (only rows of the same leaf can
// collide; leaves of one phase are distance-2 independent).
__global__ void __launch_bounds__(kBlock) k_symmspmv_chunks(
    const int* __restrict__ rp, const int* __restrict__ col, const double* __restrict__ val,
    const double* __restrict__ x, double* b, int n, int nrhs, const int* __restrict__ chunkLo,
    const int* __restrict__ chunkHi, int chunk0) {
    const int c = chunk0 + blockIdx.x;
    const int row = chunkLo[c] + threadIdx.x;
    if (row >= chunkHi[c]) return;
    const int s = rp[row], e = rp[row + 1];
    for (int k = 0; k < nrhs; ++k) {
        const size_t off = (size_t)k * n;
        const double xr = x[off + row];
        double acc = val[s] * xr;
        for (int j = s + 1; j < e; ++j) {
            const double m = val[j];
            const int cc = col[j];
            acc += m * x[off + cc];
            atomicAdd(&b[off + cc], m * xr);
        }
        atomicAdd(&b[off + row], acc);
    }
}

// b = A x over all rows
template <int TPR>
__global__ void __launch_bounds__(kBlock) k_spmv(const int* __restrict__ rp,
                                                 const int* __restrict__ col,
                                                 const double* __restrict__ val,
                                                 const double* __restrict__ x, double* __restrict__ b,
                                                 int n, int nrhs, int rowLo, int rowHi, long xoff,
                                                 long boff) {
    const int lane = threadIdx.x % TPR;
    const int row = rowLo + (blockIdx.x * blockDim.x + threadIdx.x) / TPR;
    if (row >= rowHi) return;
    const int s = rp[row], e = rp[row + 1];
    for (int k = 0; k < nrhs; ++k) {
        const size_t off = (size_t)k * n;
        double t = 0.0;
        for (int j = s + lane; j < e; j += TPR) t += val[j] * x[xoff + off + col[j]];
        t = groupSum<TPR>(t);
        if (lane == 0) b[boff + off + row] = t;
    }
}


// ---------------------------------------------------------------------------
// leaf-bundle kernels: one workgroup per bundle, TPR lanes per row, each lane
// holding E (column, value) pairs of its row in registers.
// ---------------------------------------------------------------------------
template <int TPR>
__device__ __forceinline__ double laneSum(double v) {
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, TPR);
    return v;
}

// Gauss-Seidel on one bundle.  Prologue: every lane loads its share of the
// row once; entries on columns outside the bundle (constant during the
// phase) are folded into acc = b - sum a x.  The waves then only gather the
// in-bundle columns from LDS; x of the bundle rows is written back last.
template <int BS, int TPR, int E, bool ONE_RHS>
__global__ void __launch_bounds__(BS) k_gs_bundle(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ waveSlot, const int* __restrict__ rowId,
    const double* __restrict__ diag, const int* __restrict__ scol, const double* __restrict__ sval,
    double* __restrict__ x, const double* __restrict__ b, int n, int nrhs, int bundle0, int ldsRows) {
    extern __shared__ double smem[];
    double* xl = smem;
    int* ws = reinterpret_cast<int*>(xl + ldsRows);

    const int bnd = bundle0 + blockIdx.x;
    const int s0 = bSlot[bnd], nb = bSlot[bnd + 1] - s0;
    const long off = bOff[bnd];
    const int wb = bLen[bnd];
    const int w0 = bWave[bnd], nw = bWave[bnd + 1] - w0;
    const int t = threadIdx.x, i = t / TPR, l = t % TPR;
    const bool active = i < nb;
    const long lanes = (long)nb * TPR;

    int cc[E];
    double vv[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const bool in = active && u < wb;
        const long e = off + (long)u * lanes + t;
        cc[u] = in ? scol[e] : 0;
        vv[u] = in ? sval[e] : 0.0;
    }
    const int row = active ? rowId[s0 + i] : 0;
    const double dg = active ? diag[s0 + i] : 1.0;
    for (int w = t; w <= nw; w += BS) ws[w] = waveSlot[w0 + w] - s0;
    __syncthreads();
    int myWave = -1;
    if (active)
        for (int w = 0; w < nw; ++w)
            if (i >= ws[w] && i < ws[w + 1]) myWave = w;

    // branch-free split of the register entries: global gathers use column
    // gi with coefficient gv, LDS gathers slot li with coefficient lv; the
    // other coefficient of each entry is zero
    int gi[E], li[E];
    double gv[E], lv[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const bool inner = cc[u] < 0;
        gi[u] = inner ? row : cc[u];
        li[u] = inner ? ~cc[u] : 0;
        gv[u] = inner ? 0.0 : vv[u];
        lv[u] = inner ? vv[u] : 0.0;
    }

    const int nr = ONE_RHS ? 1 : nrhs;
    for (int k = 0; k < nr; ++k) {
        const size_t vo = (size_t)k * n;
        double xv[E];
#pragma unroll
        for (int u = 0; u < E; ++u) xv[u] = x[vo + gi[u]];
        double acc = (active && l == 0) ? b[vo + row] : 0.0;
        const double xr = active ? x[vo + row] : 0.0;
#pragma unroll
        for (int u = 0; u < E; ++u) acc -= gv[u] * xv[u];
        acc = laneSum<TPR>(acc);
        if (active && l == 0) xl[i] = xr;
        __syncthreads();
        for (int w = 0; w < nw; ++w) {
            if (myWave == w) {
                double s = 0.0;
#pragma unroll
                for (int u = 0; u < E; ++u) s += lv[u] * xl[li[u]];
                s = laneSum<TPR>(s);
                if (l == 0) xl[i] = (acc - s) / dg;
            }
            __syncthreads();
        }
        if (active && l == 0) x[vo + row] = xl[i];
        __syncthreads();
    }
}

// Kaczmarz on one bundle: every column the bundle touches lives in LDS; the
// row entries stay in registers for all right-hand sides.  Rows of one wave
// share no column, so their LDS updates never collide.
template <int BS, int TPR, int E>
__global__ void __launch_bounds__(BS) k_kacz_bundle(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ waveSlot, const int* __restrict__ rowId,
    const double* __restrict__ rowNorm, const int* __restrict__ colSet,
    const int* __restrict__ colSetPtr, const int* __restrict__ scol, const double* __restrict__ sval,
    double* __restrict__ x, const double* __restrict__ b, int n, int nrhs, int bundle0, int ldsCols) {
    extern __shared__ double smem[];
    double* xl = smem;
    int* ws = reinterpret_cast<int*>(xl + ldsCols);

    const int bnd = bundle0 + blockIdx.x;
    const int s0 = bSlot[bnd], nb = bSlot[bnd + 1] - s0;
    const long off = bOff[bnd];
    const int wb = bLen[bnd];
    const int w0 = bWave[bnd], nw = bWave[bnd + 1] - w0;
    const int c0 = colSetPtr[bnd], nc = colSetPtr[bnd + 1] - c0;
    const int t = threadIdx.x, i = t / TPR, l = t % TPR;
    const bool active = i < nb;
    const long lanes = (long)nb * TPR;

    int cc[E];
    double vv[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const bool in = active && u < wb;
        const long e = off + (long)u * lanes + t;
        cc[u] = in ? ~scol[e] : 0;
        vv[u] = in ? sval[e] : 0.0;
    }
    const int row = active ? rowId[s0 + i] : 0;
    const double rn = active ? rowNorm[s0 + i] : 1.0;
    for (int w = t; w <= nw; w += BS) ws[w] = waveSlot[w0 + w] - s0;
    __syncthreads();
    int myWave = -1;
    if (active)
        for (int w = 0; w < nw; ++w)
            if (i >= ws[w] && i < ws[w + 1]) myWave = w;

    for (int k = 0; k < nrhs; ++k) {
        const size_t vo = (size_t)k * n;
        for (int c = t; c < nc; c += BS) xl[c] = x[vo + colSet[c0 + c]];
        const double br = active ? b[vo + row] : 0.0;
        __syncthreads();
        for (int w = 0; w < nw; ++w) {
            if (myWave == w) {
                double dot = 0.0;
#pragma unroll
                for (int u = 0; u < E; ++u) dot += vv[u] * xl[cc[u]];
                dot = laneSum<TPR>(dot);
                const double scale = (dot - br) / rn;
#pragma unroll
                for (int u = 0; u < E; ++u)
                    if (vv[u] != 0.0) xl[cc[u]] -= scale * vv[u];
            }
            __syncthreads();
        }
        for (int c = t; c < nc; c += BS) x[vo + colSet[c0 + c]] = xl[c];
        __syncthreads();
    }
}

inline unsigned gridFor(long n, int block = kBlock) {
    long g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

template <typename T>
int upload(DevBuf<T>& d, const T* h, size_t n) {
    RACE_HIP(d.reserve(n));
    if (n) RACE_HIP(hipMemcpy(d.p, h, n * sizeof(T), hipMemcpyHostToDevice));
    return OK;
}

// device CRS copy of a caller matrix
struct DevCrs {
    int n = 0;
    long nnz = 0;
    int maxRowLen = 0;
    DevBuf<int> rp, col;
    DevBuf<double> val;
    int load(const race_kernel_desc& d) {
        n = d.nrow;
        nnz = d.rowPtr[n];
        for (int i = 0; i < n; ++i) maxRowLen = std::max(maxRowLen, d.rowPtr[i + 1] - d.rowPtr[i]);
        int e;
        if ((e = upload(rp, d.rowPtr, (size_t)n + 1))) return e;
        if ((e = upload(col, d.col, (size_t)nnz))) return e;
        if ((e = upload(val, d.val, (size_t)nnz))) return e;
        return OK;
    }
};

struct DevWaves {
    WaveTable h;
    DevBuf<int> leafWave, wavePtr, rows;
    int load(WaveTable&& t) {
        h = std::move(t);
        int e;
        if ((e = upload(leafWave, h.leafWavePtr.data(), h.leafWavePtr.size()))) return e;
        if ((e = upload(wavePtr, h.wavePtr.data(), h.wavePtr.size()))) return e;
        if ((e = upload(rows, h.rows.data(), h.rows.size()))) return e;
        return OK;
    }
};

// threads per row from the mean row length
int pickTpr(long nnz, int n) {
    double m = n ? (double)nnz / n : 1;
    if (m <= 6) return 4;
    if (m <= 12) return 8;
    if (m <= 24) return 16;
    return 32;
}

// ---------------------------------------------------------------------------
class WaveFn : public DeviceFn {
  public:
    DevCrs A;
    DevWaves wf, wr;
    int nrhs = 1, tpr = 8;
    int init(const race_kernel_desc& d, const ExecContext& ctx) {
        kind = d.kind;
        nrhs = d.nrhs > 0 ? d.nrhs : 1;
        x = d.x;
        b = d.b;
        int e;
        if ((e = A.load(d))) return e;
        tpr = pickTpr(A.nnz, A.n);
        if ((e = wf.load(buildWaves(*ctx.fwd, kind, d.nrow, d.rowPtr, d.col)))) return e;
        if ((e = wr.load(buildWaves(*ctx.rev, kind, d.nrow, d.rowPtr, d.col)))) return e;
        return OK;
    }
    int waves(bool rev) const override { return (int)(rev ? wr.h.critical : wf.h.critical); }

    template <int TPR>
    void launchPhase(const DevWaves& w, int l0, int nl, hipStream_t st) {
        if (kind == RACE_K_GS)
            hipLaunchKernelGGL(k_gs_leaf<TPR>, dim3(nl), dim3(kBlock), 0, st, A.rp.p, A.col.p,
                               A.val.p, (double*)x, (const double*)b, A.n, nrhs, w.leafWave.p,
                               w.wavePtr.p, w.rows.p, l0);
        else
            hipLaunchKernelGGL(k_kacz_leaf<TPR>, dim3(nl), dim3(kBlock), 0, st, A.rp.p, A.col.p,
                               A.val.p, (double*)x, (const double*)b, A.n, nrhs, w.leafWave.p,
                               w.wavePtr.p, w.rows.p, l0);
    }

    int run(bool rev, hipStream_t st) override {
        const DevWaves& w = rev ? wr : wf;
        const std::vector<int>& pp = w.h.phaseLeafPtr;
        for (size_t p = 0; p + 1 < pp.size(); ++p) {
            const int l0 = pp[p], nl = pp[p + 1] - pp[p];
            if (nl <= 0) continue;
            switch (tpr) {
                case 4: launchPhase<4>(w, l0, nl, st); break;
                case 8: launchPhase<8>(w, l0, nl, st); break;
                case 16: launchPhase<16>(w, l0, nl, st); break;
                default: launchPhase<32>(w, l0, nl, st); break;
            }
        }
        RACE_HIP(hipGetLastError());
        return OK;
    }
};


// Gauss-Seidel / Kaczmarz over leaf bundles; falls back to WaveFn when a
// leaf does not fit the bundle limits.
class BundleFn : public DeviceFn {
  public:
    struct Dev {
        BundleTable h;
        DevBuf<int> slot, wave, len, waveSlot, rowId, scol, colSet, colSetPtr;
        DevBuf<long> off;
        DevBuf<double> diag, rowNorm, sval;
        int load(BundleTable&& t) {
            h = std::move(t);
            int e;
            if ((e = upload(slot, h.slot.data(), h.slot.size()))) return e;
            if ((e = upload(wave, h.wave.data(), h.wave.size()))) return e;
            if ((e = upload(len, h.len.data(), h.len.size()))) return e;
            if ((e = upload(off, h.off.data(), h.off.size()))) return e;
            if ((e = upload(waveSlot, h.waveSlot.data(), h.waveSlot.size()))) return e;
            if ((e = upload(rowId, h.rowId.data(), h.rowId.size()))) return e;
            if ((e = upload(scol, h.scol.data(), h.scol.size()))) return e;
            if ((e = upload(sval, h.sval.data(), h.sval.size()))) return e;
            if (h.kind == RACE_K_KACZ) {
                if ((e = upload(rowNorm, h.rowNorm.data(), h.rowNorm.size()))) return e;
                if ((e = upload(colSet, h.colSet.data(), h.colSet.size()))) return e;
                if ((e = upload(colSetPtr, h.colSetPtr.data(), h.colSetPtr.size()))) return e;
            } else {
                if ((e = upload(diag, h.diag.data(), h.diag.size()))) return e;
            }
            return OK;
        }
        size_t lds(int p) const {
            const int held = h.kind == RACE_K_KACZ ? h.phCols[p] : h.phRows[p];
            return (size_t)held * sizeof(double) + (size_t)(h.phWaves[p] + 1) * sizeof(int) + 16;
        }
    };
    Dev df, dr;
    int nrhs = 1, n = 0, bs = 256, tpr = 2, E = 16;
    std::unique_ptr<WaveFn> fallback;

    // register widths with a kernel instantiation, per lanes-per-row
    static int pickE(int tpr, int need) {
        static const int e1[] = {4, 8, 16, 32}, e2[] = {4, 8, 13, 16}, e4[] = {2, 4, 7, 8};
        const int* c = tpr == 1 ? e1 : tpr == 2 ? e2 : e4;
        for (int q = 0; q < 4; ++q)
            if (c[q] >= need) return c[q];
        return -1;
    }

    int init(const race_kernel_desc& d, const ExecContext& ctx, int cap, int innerCap) {
        kind = d.kind;
        nrhs = d.nrhs > 0 ? d.nrhs : 1;
        n = d.nrow;
        x = d.x;
        b = d.b;
        bs = ctx.bundleThreads;
        int maxLen = 0;
        for (int r = 0; r < n; ++r) maxLen = std::max(maxLen, d.rowPtr[r + 1] - d.rowPtr[r]);
        if (kind != RACE_K_KACZ) maxLen -= 1;  // the diagonal is kept apart
        tpr = ctx.bundleTpr > 0 ? ctx.bundleTpr : (maxLen <= 8 ? 1 : maxLen <= 32 ? 2 : 4);
        E = pickE(tpr, (maxLen + tpr - 1) / tpr);
        const int rowsCap = std::min(cap, bs / tpr);
        if (E < 0 || rowsCap < 1) {
            if (ctx.layout == 2) return ERR_INCOMPATIBILITY;
            fallback = std::make_unique<WaveFn>();
            return fallback->init(d, ctx);
        }
        BundleTable tf = buildBundles(*ctx.fwd, kind, d.nrow, d.rowPtr, d.col, d.val, rowsCap, innerCap, tpr);
        BundleTable tr = buildBundles(*ctx.rev, kind, d.nrow, d.rowPtr, d.col, d.val, rowsCap, innerCap, tpr);
        if (!tf.ok || !tr.ok) {
            if (ctx.layout == 2) return ERR_INCOMPATIBILITY;  // bundles required
            fallback = std::make_unique<WaveFn>();
            return fallback->init(d, ctx);
        }
        int e;
        if ((e = df.load(std::move(tf)))) return e;
        if ((e = dr.load(std::move(tr)))) return e;
        for (int p = 0; p < df.h.numPhases(); ++p)
            if (df.lds(p) > kMaxLds) return ERR_INTERNAL;
        for (int p = 0; p < dr.h.numPhases(); ++p)
            if (dr.lds(p) > kMaxLds) return ERR_INTERNAL;
        return OK;
    }

    template <int BS, int TPR, int EE>
    void launch(const Dev& d, int b0, int nbnd, int p, hipStream_t st) {
        if (kind == RACE_K_KACZ)
            hipLaunchKernelGGL((k_kacz_bundle<BS, TPR, EE>), dim3(nbnd), dim3(BS), d.lds(p), st, d.slot.p,
                               d.wave.p, d.off.p, d.len.p, d.waveSlot.p, d.rowId.p, d.rowNorm.p,
                               d.colSet.p, d.colSetPtr.p, d.scol.p, d.sval.p, (double*)x,
                               (const double*)b, n, nrhs, b0, d.h.phCols[p]);
        else if (nrhs == 1)
            hipLaunchKernelGGL((k_gs_bundle<BS, TPR, EE, true>), dim3(nbnd), dim3(BS), d.lds(p), st,
                               d.slot.p, d.wave.p, d.off.p, d.len.p, d.waveSlot.p, d.rowId.p, d.diag.p,
                               d.scol.p, d.sval.p, (double*)x, (const double*)b, n, nrhs, b0,
                               d.h.phRows[p]);
        else
            hipLaunchKernelGGL((k_gs_bundle<BS, TPR, EE, false>), dim3(nbnd), dim3(BS), d.lds(p), st,
                               d.slot.p, d.wave.p, d.off.p, d.len.p, d.waveSlot.p, d.rowId.p, d.diag.p,
                               d.scol.p, d.sval.p, (double*)x, (const double*)b, n, nrhs, b0,
                               d.h.phRows[p]);
    }
    template <int BS>
    void launchBs(const Dev& d, int b0, int nbnd, int p, hipStream_t st) {
        switch (tpr * 100 + E) {
            case 104: launch<BS, 1, 4>(d, b0, nbnd, p, st); break;
            case 108: launch<BS, 1, 8>(d, b0, nbnd, p, st); break;
            case 116: launch<BS, 1, 16>(d, b0, nbnd, p, st); break;
            case 132: launch<BS, 1, 32>(d, b0, nbnd, p, st); break;
            case 204: launch<BS, 2, 4>(d, b0, nbnd, p, st); break;
            case 208: launch<BS, 2, 8>(d, b0, nbnd, p, st); break;
            case 213: launch<BS, 2, 13>(d, b0, nbnd, p, st); break;
            case 216: launch<BS, 2, 16>(d, b0, nbnd, p, st); break;
            case 402: launch<BS, 4, 2>(d, b0, nbnd, p, st); break;
            case 404: launch<BS, 4, 4>(d, b0, nbnd, p, st); break;
            case 407: launch<BS, 4, 7>(d, b0, nbnd, p, st); break;
            default: launch<BS, 4, 8>(d, b0, nbnd, p, st); break;
        }
    }
    int waves(bool rev) const override {
        if (fallback) return fallback->waves(rev);
        return (int)(rev ? dr.h.critical : df.h.critical);
    }
    int run(bool rev, hipStream_t st) override {
        if (fallback) {
            fallback->x = x;
            fallback->b = b;
            return fallback->run(rev, st);
        }
        const Dev& d = rev ? dr : df;
        const std::vector<int>& pb = d.h.phaseBundle;
        for (int p = 0; p + 1 < (int)pb.size(); ++p) {
            const int b0 = pb[p], nbnd = pb[p + 1] - pb[p];
            if (nbnd <= 0) continue;
            switch (bs) {
                case 64: launchBs<64>(d, b0, nbnd, p, st); break;
                case 128: launchBs<128>(d, b0, nbnd, p, st); break;
                case 512: launchBs<512>(d, b0, nbnd, p, st); break;
                case 1024: launchBs<1024>(d, b0, nbnd, p, st); break;
                default: launchBs<256>(d, b0, nbnd, p, st); break;
            }
        }
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

class SymmSpmvFn : public DeviceFn {
  public:
    DevCrs A;
    int nrhs = 1;
    std::vector<int> phaseChunk, hLo, hHi;
    DevBuf<int> cLo, cHi;
    int init(const race_kernel_desc& d, const ExecContext& ctx) {
        kind = d.kind;
        nrhs = d.nrhs > 0 ? d.nrhs : 1;
        x = d.x;
        b = d.b;
        int e;
        if ((e = A.load(d))) return e;
        const Schedule& s = *ctx.fwd;
        phaseChunk.push_back(0);
        for (int p = 0; p < s.numPhases(); ++p) {
            for (int l = s.phasePtr[p]; l < s.phasePtr[p + 1]; ++l)
                for (int r = s.leafLo[l]; r < s.leafHi[l]; r += kBlock) {
                    hLo.push_back(r);
                    hHi.push_back(std::min(r + kBlock, s.leafHi[l]));
                }
            phaseChunk.push_back((int)hLo.size());
        }
        if ((e = upload(cLo, hLo.data(), hLo.size()))) return e;
        if ((e = upload(cHi, hHi.data(), hHi.size()))) return e;
        return OK;
    }
    int run(bool, hipStream_t st) override {
        for (size_t p = 0; p + 1 < phaseChunk.size(); ++p) {
            const int c0 = phaseChunk[p], nc = phaseChunk[p + 1] - phaseChunk[p];
            if (nc <= 0) continue;
            hipLaunchKernelGGL(k_symmspmv_chunks, dim3(nc), dim3(kBlock), 0, st, A.rp.p, A.col.p,
                               A.val.p, (const double*)x, (double*)b, A.n, nrhs, cLo.p, cHi.p, c0);
        }
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

class SpmvFn : public DeviceFn {
  public:
    DevCrs A;
    int nrhs = 1, tpr = 8;
    int init(const race_kernel_desc& d) {
        kind = d.kind;
        nrhs = d.nrhs > 0 ? d.nrhs : 1;
        x = d.x;
        b = d.b;
        int e;
        if ((e = A.load(d))) return e;
        tpr = pickTpr(A.nnz, A.n);
        return OK;
    }
    template <int TPR>
    void launch(int lo, int hi, long xo, long bo, double* xx, double* bb, hipStream_t st) {
        const long threads = (long)(hi - lo) * TPR;
        hipLaunchKernelGGL(k_spmv<TPR>, dim3(gridFor(threads)), dim3(kBlock), 0, st, A.rp.p,
                           A.col.p, A.val.p, xx, bb, A.n, nrhs, lo, hi, xo, bo);
    }
    void rows(int lo, int hi, long xo, long bo, double* xx, double* bb, hipStream_t st) {
        if (hi <= lo) return;
        switch (tpr) {
            case 4: launch<4>(lo, hi, xo, bo, xx, bb, st); break;
            case 8: launch<8>(lo, hi, xo, bo, xx, bb, st); break;
            case 16: launch<16>(lo, hi, xo, bo, xx, bb, st); break;
            default: launch<32>(lo, hi, xo, bo, xx, bb, st); break;
        }
    }
    int run(bool, hipStream_t st) override {
        rows(0, A.n, 0, 0, (double*)x, (double*)b, st);
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

// Level-blocked matrix powers: level group g at power p runs after groups
// g-1, g, g+1 at power p-1 (BFS levels only touch neighbouring levels).  The
// launch order is the reference's trapezoidal wavefront: step L runs powers
// 1..k on groups L, L-1, ..., L-k+1 (src/functionManager.cpp:618-822).
class MpkFn : public SpmvFn {
  public:
    std::vector<int> groupPtr;
    int init(const race_kernel_desc& d, const ExecContext& ctx) {
        int e = SpmvFn::init(d);
        if (e) return e;
        if (!ctx.mpk) return ERR_INVALID_ARG;
        groupPtr = ctx.mpk->groupPtr;
        power = d.power > 0 ? d.power : 1;
        nrhs = 1;
        return OK;
    }
    int run(bool, hipStream_t st) override {
        const int G = (int)groupPtr.size() - 1;
        const long n = A.n;
        double* xx = (double*)x;
        for (int L = 0; L < G + power - 1; ++L)
            for (int p = 0; p < power; ++p) {
                const int g = L - p;
                if (g < 0 || g >= G) continue;
                rows(groupPtr[g], groupPtr[g + 1], (long)p * n, (long)(p + 1) * n, xx, xx, st);
            }
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

}  // namespace

namespace {
// plain streaming copy (16-byte vectors), used as the byte-counter reference
__global__ void __launch_bounds__(kBlock) k_copy(const int4* __restrict__ src,
                                                  int4* __restrict__ dst, long n16) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}
}  // namespace

int deviceCopy(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes % 16) return ERR_INVALID_ARG;
    const long n16 = (long)(bytes / 16);
    hipLaunchKernelGGL(k_copy, dim3(1024 * 8), dim3(kBlock), 0, st, (const int4*)src, (int4*)dst, n16);
    RACE_HIP(hipGetLastError());
    return OK;
}

int makeDeviceFn(const race_kernel_desc& d, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out) {
    if (d.nrow <= 0 || !d.rowPtr || !d.col || !d.val) return ERR_INVALID_ARG;
    if (d.val_precision != RACE_FP64 || d.vec_precision != RACE_FP64) return ERR_NOT_IMPLEMENTED;
    int e = OK;
    switch (d.kind) {
        case RACE_K_GS:
        case RACE_K_KACZ: {
            if (!ctx.fwd || !ctx.rev) return ERR_INVALID_ARG;
            if (ctx.layout == 1) {
                auto f = std::make_unique<WaveFn>();
                e = f->init(d, ctx);
                *out = std::move(f);
            } else {
                auto f = std::make_unique<BundleFn>();
                e = f->init(d, ctx, ctx.bundleRows, ctx.bundleInner);
                *out = std::move(f);
            }
            break;
        }
        case RACE_K_SYMMSPMV: {
            if (!ctx.fwd) return ERR_INVALID_ARG;
            auto f = std::make_unique<SymmSpmvFn>();
            e = f->init(d, ctx);
            *out = std::move(f);
            break;
        }
        case RACE_K_SPMV: {
            auto f = std::make_unique<SpmvFn>();
            e = f->init(d);
            *out = std::move(f);
            break;
        }
        case RACE_K_MPK: {
            auto f = std::make_unique<MpkFn>();
            e = f->init(d, ctx);
            *out = std::move(f);
            break;
        }
        default:
            return ERR_INVALID_ARG;
    }
    return e;
}

}  // namespace race
