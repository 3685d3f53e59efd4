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
#include <cstdint>
#include <cstring>
#include <type_traits>

#include "bundle.hpp"
#include "exec.hpp"
#include "exec_common.hpp"

namespace race {

namespace {

// matrix value storage: double, float, or bfloat16 bits (uint16_t)
template <typename XT>
__device__ __forceinline__ XT loadVal(const double* p) {
    return (XT)*p;
}
template <typename XT>
__device__ __forceinline__ XT loadVal(const float* p) {
    return (XT)*p;
}
template <typename XT>
__device__ __forceinline__ XT loadVal(const uint16_t* p) {
    return (XT)__uint_as_float((unsigned)*p << 16);
}

template <int TPR, typename T>
__device__ __forceinline__ T laneSum(T v) {
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, TPR);
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

template <int BS, int TPR, int E, typename VT, typename XT, bool ONE_RHS>
__global__ void __launch_bounds__(BS) k_gs_bundle(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const int* __restrict__ waveSlot,
    const int* __restrict__ rowId, const XT* __restrict__ diag, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, XT* __restrict__ x,
    const XT* __restrict__ b, int n, int nrhs, int bundle0, int nbnd, int xcd, int ldsLocal) {
    extern __shared__ __align__(16) unsigned char smem[];
    XT* xl = reinterpret_cast<XT*>(smem);
    int* ws = reinterpret_cast<int*>(xl + ldsLocal);

    const int q = bundleOfBlock(nbnd, xcd);
    if (q < 0) return;
    const int bnd = bundle0 + q;
    const int s0 = bSlot[bnd], nb = bSlot[bnd + 1] - s0;
    const long off = bOff[bnd];
    const int wb = bLen[bnd];
    const int w0 = bWave[bnd], nw = bWave[bnd + 1] - w0;
    const int e0 = bExt[bnd], ne = bExt[bnd + 1] - e0;
    const int t = threadIdx.x, i = t / TPR, l = t % TPR;
    const bool active = i < nb;
    const long lanes = (long)nb * TPR;

    int li[E];
    XT lv[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const bool in = active && u < wb;
        const long e = off + (long)u * lanes + t;
        li[u] = in ? (int)sidx[e] : 0;
        lv[u] = in ? loadVal<XT>(sval + e) : XT(0);
    }
    const int row = active ? rowId[s0 + i] : 0;
    const XT dg = active ? diag[s0 + i] : XT(1);
    for (int w = t; w <= nw; w += BS) ws[w] = waveSlot[w0 + w] - s0;
    __syncthreads();
    int myWave = -1;
    if (active)
        for (int w = 0; w < nw; ++w)
            if (i >= ws[w] && i < ws[w + 1]) myWave = w;

    const int nr = ONE_RHS ? 1 : nrhs;
    for (int k = 0; k < nr; ++k) {
        const size_t vo = (size_t)k * n;
        for (int c = t; c < ne; c += BS) xl[nb + c] = x[vo + ext[e0 + c]];
        if (active && l == 0) xl[i] = x[vo + row];
        const XT br = (active && l == 0) ? b[vo + row] : XT(0);
        __syncthreads();
        for (int w = 0; w < nw; ++w) {
            if (myWave == w) {
                XT s = XT(0);
#pragma unroll
                for (int u = 0; u < E; ++u) s += lv[u] * xl[li[u]];
                s = laneSum<TPR>(s);
                if (l == 0) xl[i] = (br - s) / dg;
            }
            __syncthreads();
        }
        if (active && l == 0) x[vo + row] = xl[i];
        __syncthreads();
    }
}

// Kaczmarz: every column the bundle touches lives in LDS; rows of one wave
// share no column.  The row norms stay FP64.
template <int BS, int TPR, int E, typename VT, typename XT>
__global__ void __launch_bounds__(BS) k_kacz_bundle(
    const int* __restrict__ bSlot, const int* __restrict__ bWave, const long* __restrict__ bOff,
    const int* __restrict__ bLen, const int* __restrict__ bExt, const int* __restrict__ waveSlot,
    const int* __restrict__ rowId, const double* __restrict__ rowNorm, const int* __restrict__ ext,
    const uint16_t* __restrict__ sidx, const VT* __restrict__ sval, XT* __restrict__ x,
    const XT* __restrict__ b, int n, int nrhs, int bundle0, int nbnd, int xcd, int ldsLocal) {
    extern __shared__ __align__(16) unsigned char smem[];
    XT* xl = reinterpret_cast<XT*>(smem);
    int* ws = reinterpret_cast<int*>(xl + ldsLocal);

    const int q = bundleOfBlock(nbnd, xcd);
    if (q < 0) return;
    const int bnd = bundle0 + q;
    const int s0 = bSlot[bnd], nb = bSlot[bnd + 1] - s0;
    const long off = bOff[bnd];
    const int wb = bLen[bnd];
    const int w0 = bWave[bnd], nw = bWave[bnd + 1] - w0;
    const int e0 = bExt[bnd], ne = bExt[bnd + 1] - e0;
    const int t = threadIdx.x, i = t / TPR, l = t % TPR;
    const bool active = i < nb;
    const long lanes = (long)nb * TPR;

    int li[E];
    XT lv[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const bool in = active && u < wb;
        const long e = off + (long)u * lanes + t;
        li[u] = in ? (int)sidx[e] : 0;
        lv[u] = in ? loadVal<XT>(sval + e) : XT(0);
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
        for (int c = t; c < ne; c += BS) xl[c] = x[vo + ext[e0 + c]];
        const XT br = active ? b[vo + row] : XT(0);
        __syncthreads();
        for (int w = 0; w < nw; ++w) {
            if (myWave == w) {
                XT dot = XT(0);
#pragma unroll
                for (int u = 0; u < E; ++u) dot += lv[u] * xl[li[u]];
                dot = laneSum<TPR>(dot);
                const XT scale = (XT)(((double)dot - (double)br) / rn);
#pragma unroll
                for (int u = 0; u < E; ++u)
                    if (lv[u] != XT(0)) xl[li[u]] -= scale * lv[u];
            }
            __syncthreads();
        }
        for (int c = t; c < ne; c += BS) x[vo + ext[e0 + c]] = xl[c];
        __syncthreads();
    }
}

uint16_t toBf16(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

// (lanes per row, register width) pairs with a kernel instantiation
int pickE(int tpr, int need) {
    static const int e1[] = {4, 8, 16}, e2[] = {8, 13, 16}, e4[] = {8};
    const int* c = tpr == 1 ? e1 : tpr == 2 ? e2 : e4;
    const int nc = tpr == 1 ? 3 : tpr == 2 ? 3 : 1;
    for (int q = 0; q < nc; ++q)
        if (c[q] >= need) return c[q];
    return -1;
}

template <typename VT, typename XT>
class BundleFn : public DeviceFn {
  public:
    struct Dev {
        BundleTable h;
        DevBuf<int> slot, wave, len, ext, extPtr, waveSlot, rowId;
        DevBuf<long> off;
        DevBuf<uint16_t> sidx;
        DevBuf<VT> sval;
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
            if ((e = upload(waveSlot, h.waveSlot.data(), h.waveSlot.size()))) return e;
            if ((e = upload(rowId, h.rowId.data(), h.rowId.size()))) return e;
            if ((e = upload(sidx, h.sidx.data(), h.sidx.size()))) return e;
            std::vector<VT> v(h.sval.size());
            for (size_t q = 0; q < v.size(); ++q) v[q] = conv(h.sval[q]);
            if ((e = upload(sval, v.data(), v.size()))) return e;
            if (h.kind == RACE_K_KACZ) {
                if ((e = upload(rowNorm, h.rowNorm.data(), h.rowNorm.size()))) return e;
            } else {
                std::vector<XT> dg(h.diag.begin(), h.diag.end());
                if ((e = upload(diag, dg.data(), dg.size()))) return e;
            }
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
        size_t lds(int p) const {
            return (size_t)h.phLocal[p] * sizeof(XT) + (size_t)(h.phWaves[p] + 1) * sizeof(int) + 32;
        }
    };
    Dev df, dr;
    int nrhs = 1, n = 0, bs = 256, tpr = 2, E = 16, xcd = 1;
    std::unique_ptr<DeviceFn> fallback;

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
        if (kind != RACE_K_KACZ) maxLen -= 1;  // the diagonal is kept apart
        maxLen = std::max(maxLen, 1);
        tpr = ctx.bundleTpr > 0 ? ctx.bundleTpr : (maxLen <= 16 ? 1 : maxLen <= 32 ? 2 : 4);
        E = pickE(tpr, (maxLen + tpr - 1) / tpr);
        const int rowsCap = std::min(ctx.bundleRows, bs / tpr);
        const bool fp64 = std::is_same<VT, double>::value && std::is_same<XT, double>::value;
        auto fallBack = [&]() -> int {
            if (ctx.layout == 2 || !fp64) return ERR_INCOMPATIBILITY;  // bundles required
            return makeWaveFn(d, ctx, &fallback);
        };
        if (E < 0 || rowsCap < 1) return fallBack();
        BundleTable tf = buildBundles(*ctx.fwd, kind, n, d.rowPtr, d.col, d.val, rowsCap,
                                      ctx.bundleLocal, tpr);
        if (!tf.ok) return fallBack();
        BundleTable tr = buildBundles(*ctx.rev, kind, n, d.rowPtr, d.col, d.val, rowsCap,
                                      ctx.bundleLocal, tpr);
        if (!tr.ok) return fallBack();
        int e;
        if ((e = df.load(std::move(tf)))) return e;
        if ((e = dr.load(std::move(tr)))) return e;
        for (int p = 0; p < df.h.numPhases(); ++p)
            if (df.lds(p) > kMaxLds) return fallBack();
        for (int p = 0; p < dr.h.numPhases(); ++p)
            if (dr.lds(p) > kMaxLds) return fallBack();
        return OK;
    }

    template <int BS, int TPR, int EE>
    void launch(const Dev& d, int b0, int nbnd, int p, hipStream_t st) {
        const unsigned grid = xcd ? 8u * (unsigned)((nbnd + 7) / 8) : (unsigned)nbnd;
        if (kind == RACE_K_KACZ)
            hipLaunchKernelGGL((k_kacz_bundle<BS, TPR, EE, VT, XT>), dim3(grid), dim3(BS), d.lds(p), st,
                               d.slot.p, d.wave.p, d.off.p, d.len.p, d.extPtr.p, d.waveSlot.p,
                               d.rowId.p, d.rowNorm.p, d.ext.p, d.sidx.p, d.sval.p, (XT*)x,
                               (const XT*)b, n, nrhs, b0, nbnd, xcd, d.h.phLocal[p]);
        else if (nrhs == 1)
            hipLaunchKernelGGL((k_gs_bundle<BS, TPR, EE, VT, XT, true>), dim3(grid), dim3(BS), d.lds(p),
                               st, d.slot.p, d.wave.p, d.off.p, d.len.p, d.extPtr.p, d.waveSlot.p,
                               d.rowId.p, d.diag.p, d.ext.p, d.sidx.p, d.sval.p, (XT*)x, (const XT*)b,
                               n, nrhs, b0, nbnd, xcd, d.h.phLocal[p]);
        else
            hipLaunchKernelGGL((k_gs_bundle<BS, TPR, EE, VT, XT, false>), dim3(grid), dim3(BS), d.lds(p),
                               st, d.slot.p, d.wave.p, d.off.p, d.len.p, d.extPtr.p, d.waveSlot.p,
                               d.rowId.p, d.diag.p, d.ext.p, d.sidx.p, d.sval.p, (XT*)x, (const XT*)b,
                               n, nrhs, b0, nbnd, xcd, d.h.phLocal[p]);
    }
    template <int BS>
    void launchBs(const Dev& d, int b0, int nbnd, int p, hipStream_t st) {
        switch (tpr * 100 + E) {
            case 104: launch<BS, 1, 4>(d, b0, nbnd, p, st); break;
            case 108: launch<BS, 1, 8>(d, b0, nbnd, p, st); break;
            case 116: launch<BS, 1, 16>(d, b0, nbnd, p, st); break;
            case 208: launch<BS, 2, 8>(d, b0, nbnd, p, st); break;
            case 213: launch<BS, 2, 13>(d, b0, nbnd, p, st); break;
            case 216: launch<BS, 2, 16>(d, b0, nbnd, p, st); break;
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

}  // namespace

int makeBundleFn(const race_kernel_desc& d, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out) {
    if (ctx.bundleThreads != 128 && ctx.bundleThreads != 256 && ctx.bundleThreads != 512 &&
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
