// exec.hip -- colour-phase launcher: SymmSpMV, Gauss-Seidel, Kaczmarz, SpMV
// and the level-blocked matrix power kernel on the MI355X.
#include <algorithm>
#include <cstring>

#include "exec.hpp"
#include "exec_common.hpp"

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

    bool graphable() const override { return true; }
    int launches(bool rev) const override {
        const std::vector<int>& pp = (rev ? wr : wf).h.phaseLeafPtr;
        int k = 0;
        for (size_t p = 0; p + 1 < pp.size(); ++p) k += pp[p + 1] > pp[p];
        return k;
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
    bool graphable() const override { return true; }
    int launches(bool) const override {
        int k = 0;
        for (size_t p = 0; p + 1 < phaseChunk.size(); ++p) k += phaseChunk[p + 1] > phaseChunk[p];
        return k;
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

}  // namespace

int makeWaveFn(const race_kernel_desc& d, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out) {
    auto f = std::make_unique<WaveFn>();
    int e = f->init(d, ctx);
    *out = std::move(f);
    return e;
}

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
    const bool fp64 = d.val_precision == RACE_FP64 && d.vec_precision == RACE_FP64;
    const bool bundleKind = d.kind == RACE_K_GS || d.kind == RACE_K_KACZ;
    // FP32 / BF16 storage: the bundle kernels and the SELL kernels (SpMV,
    // SpMTV, SymmSpMV over the expanded matrix, MPK); the CRS kernels are FP64
    const bool sellKind = d.kind == RACE_K_SPMV || d.kind == RACE_K_SPMTV || d.kind == RACE_K_MPK ||
                          d.kind == RACE_K_MPK_HIGHEST || (d.kind == RACE_K_SYMMSPMV && ctx.symmMode != 1);
    if (!fp64 && !(bundleKind && ctx.layout != 1) && !sellKind) return ERR_NOT_IMPLEMENTED;
    if (d.vec_precision != RACE_FP64 && d.vec_precision != RACE_FP32) return ERR_INVALID_ARG;
    if (d.val_precision < RACE_FP64 || d.val_precision > RACE_BF16) return ERR_INVALID_ARG;
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
                e = makeBundleFn(d, ctx, out);
            }
            break;
        }
        case RACE_K_SYMMSPMV: {
            if (!ctx.fwd) return ERR_INVALID_ARG;
            if (ctx.symmMode != 1) {
                // default: the expanded matrix (one launch, no atomics; measured
                // fastest on 7-point 64^3: 5.5 us against 11.2 us with atomics)
                e = makeSymmSellFn(d, ctx.symmMode == 2 ? 2 : 3, out);
                break;
            }
            auto f = std::make_unique<SymmSpmvFn>();
            e = f->init(d, ctx);
            *out = std::move(f);
            break;
        }
        case RACE_K_SPMV: {
            if (d.nrhs <= 1 || !fp64) {
                e = makeSellSpmvFn(d, out);
                break;
            }
            auto f = std::make_unique<SpmvFn>();
            e = f->init(d);
            *out = std::move(f);
            break;
        }
        case RACE_K_SPMTV:
            e = makeSpmtvFn(d, out);
            break;
        case RACE_K_MPK:
        case RACE_K_MPK_HIGHEST:
            e = makeMpkFn(d, ctx, out);
            break;
        default:
            return ERR_INVALID_ARG;
    }
    if (*out) {
        (*out)->vecPrecision = d.vec_precision;
        (*out)->nrow = d.nrow;
        (*out)->nrhs = d.nrhs > 0 ? d.nrhs : 1;
    }
    return e;
}

}  // namespace race
