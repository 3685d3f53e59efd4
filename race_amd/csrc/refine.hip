// refine.hip -- FP64 iterative refinement around a low-precision sweep.
//
// One refinement step: r = b - A x in FP64 over the SELL layout (and ||r||,
// ||b|| by reduction),
// d = 0, `sweeps` forward+backward sweeps of the inner Gauss-Seidel or
// Kaczmarz function on A d = r in the inner function's precision, x += d in
// FP64.  FP64 is used only for the residual and the update (and the
// Kaczmarz row norms inside the inner function).  With an FP64 inner
// function the step is exactly `sweeps` plain symmetric sweeps on A x = b
// (the sweep is affine, so the correction form gives the same iterate),
// which is how the plain FP64 path is run for comparison.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>

#include "exec.hpp"
#include "exec_common.hpp"

namespace race {

namespace {

template <typename T>
__global__ void k_zero(T* __restrict__ v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) v[i] = T(0);
}

template <typename T>
__global__ void k_update(double* __restrict__ x, const T* __restrict__ d, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) x[i] += (double)d[i];
}

// Device workspace of one refinement: the residual r and the correction d
// (n * nrhs entries each, in the inner function's vector precision) and the
// two squared norms.
struct RefineWork {
    double* norms = nullptr;
    void* r = nullptr;
    void* d = nullptr;
};

class RefineFn : public DeviceFn {
  public:
    int n = 0, nrhs = 1, sweeps = 1;
    std::unique_ptr<ResidualOp> res;
    DevBuf<double> norms, r64, d64;
    DevBuf<float> r32, d32;
    DeviceFn* inner = nullptr;
    double* hostNorms = nullptr;
    double lastRelres = -1;
    long totalSweeps = 0;

    ~RefineFn() override {
        if (hostNorms) (void)hipHostFree(hostNorms);
    }

    bool innerFp32() const { return inner->vecPrecision == RACE_FP32; }
    size_t vecBytes() const { return (size_t)n * nrhs * (innerFp32() ? sizeof(float) : sizeof(double)); }
    RefineWork own() {
        RefineWork w;
        w.norms = norms.p;
        w.r = innerFp32() ? (void*)r32.p : (void*)r64.p;
        w.d = innerFp32() ? (void*)d32.p : (void*)d64.p;
        return w;
    }

    int init(const race_kernel_desc& d, DeviceFn* in, int innerSweeps) {
        if (!in || (in->kind != RACE_K_GS && in->kind != RACE_K_KACZ) || innerSweeps < 1)
            return ERR_INVALID_ARG;
        if (d.val_precision != RACE_FP64 || d.vec_precision != RACE_FP64) return ERR_INVALID_ARG;
        // the inner sweeps run on this function's workspace: same rows and
        // right-hand sides, or they would index past it
        if (in->nrow != d.nrow || in->nrhs != (d.nrhs > 0 ? d.nrhs : 1)) return ERR_INVALID_ARG;
        kind = d.kind;
        inner = in;
        sweeps = innerSweeps;
        n = d.nrow;
        nrhs = d.nrhs > 0 ? d.nrhs : 1;
        x = d.x;
        b = d.b;
        int e;
        if ((e = makeSellResidual(d, &res))) return e;
        RACE_HIP(norms.reserve(2));
        const size_t len = (size_t)n * nrhs;
        if (innerFp32()) {
            RACE_HIP(r32.reserve(len));
            RACE_HIP(d32.reserve(len));
        } else {
            RACE_HIP(r64.reserve(len));
            RACE_HIP(d64.reserve(len));
        }
        RACE_HIP(hipHostMalloc(reinterpret_cast<void**>(&hostNorms), 2 * sizeof(double)));
        return OK;
    }

    // r = b - A x; relres = ||r|| / ||b|| when requested (synchronises)
    int residual(const RefineWork& w, hipStream_t st, double* relres) {
        hipLaunchKernelGGL(k_zero<double>, dim3(1), dim3(64), 0, st, w.norms, 2l);
        int e = res->run((const double*)x, (const double*)b, w.r, innerFp32(), nrhs, w.norms, st);
        if (e) return e;
        if (relres) {
            RACE_HIP(hipMemcpyAsync(hostNorms, w.norms, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
            RACE_HIP(hipStreamSynchronize(st));
            *relres = hostNorms[1] > 0 ? std::sqrt(hostNorms[0] / hostNorms[1]) : std::sqrt(hostNorms[0]);
            lastRelres = *relres;
        }
        return OK;
    }

    // correction sweeps on A d = r and x += d (residual already in r).  The
    // inner function is pointed at the workspace for the sweeps and back at
    // its own vectors afterwards, so it stays usable on its own.
    int correct(const RefineWork& w, hipStream_t st) {
        const long len = (long)n * nrhs;
        const unsigned g = gridFor(len);
        const bool f32 = innerFp32();
        if (f32)
            hipLaunchKernelGGL(k_zero<float>, dim3(g), dim3(kBlock), 0, st, (float*)w.d, len);
        else
            hipLaunchKernelGGL(k_zero<double>, dim3(g), dim3(kBlock), 0, st, (double*)w.d, len);
        void* const ownX = inner->x;
        void* const ownB = inner->b;
        inner->x = w.d;
        inner->b = w.r;
        int e = OK;
        for (int s = 0; s < sweeps && !e; ++s) {
            e = inner->run(false, st);
            if (!e) e = inner->run(true, st);
        }
        inner->x = ownX;
        inner->b = ownB;
        if (e) return e;
        totalSweeps += sweeps;
        if (f32)
            hipLaunchKernelGGL(k_update<float>, dim3(g), dim3(kBlock), 0, st, (double*)x, (const float*)w.d, len);
        else
            hipLaunchKernelGGL(k_update<double>, dim3(g), dim3(kBlock), 0, st, (double*)x, (const double*)w.d,
                               len);
        RACE_HIP(hipGetLastError());
        return OK;
    }

    // one refinement step (asynchronous)
    int run(bool, hipStream_t st) override {
        const RefineWork w = own();
        int e = residual(w, st, nullptr);
        if (e) return e;
        return correct(w, st);
    }

    int solve(const RefineWork& w, double tol, int maxOuter, hipStream_t st, int* outer, double* relres) {
        double rr = 0;
        int it = 0;
        int e = residual(w, st, &rr);
        if (e) return e;
        while (rr > tol && it < maxOuter) {
            if ((e = correct(w, st))) return e;
            ++it;
            if ((e = residual(w, st, &rr))) return e;
        }
        if (outer) *outer = it;
        if (relres) *relres = rr;
        return OK;
    }
};

}  // namespace

int makeRefineFn(const race_kernel_desc& d, DeviceFn* inner, int sweeps, std::unique_ptr<DeviceFn>* out) {
    auto f = std::make_unique<RefineFn>();
    int e = f->init(d, inner, sweeps);
    *out = std::move(f);
    return e;
}

int refineSolve(DeviceFn* f, double tol, int maxOuter, hipStream_t st, int* outer, long* sweeps,
                double* relres) {
    RefineFn* r = dynamic_cast<RefineFn*>(f);
    if (!r) return ERR_INVALID_ARG;
    const long before = r->totalSweeps;
    int e = r->solve(r->own(), tol, maxOuter, st, outer, relres);
    if (sweeps) *sweeps = r->totalSweeps - before;
    return e;
}

int refineResidual(DeviceFn* f, hipStream_t st, double* relres) {
    RefineFn* r = dynamic_cast<RefineFn*>(f);
    if (!r) return ERR_INVALID_ARG;
    return r->residual(r->own(), st, relres);
}

// ---------------------------------------------------------------------------
// The refinement without allocation.  Its workspace (norms, r, d) lives in
// one static device array of kScratchBytes; every GPU the library runs on
// holds its own copy until the process ends, and calls on one GPU share it,
// so concurrent calls on one GPU race.  Host code never uses the array's
// name as an address: the device address comes from hipGetSymbolAddress,
// once per device, with that device current.
namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchAlign = 256;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex scratchLock;
void* scratchPtr[kMaxDevices] = {};

int scratchOfCurrentDevice(void** out) {
    int dev = -1;
    RACE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return ERR_INTERNAL;
    std::lock_guard<std::mutex> g(scratchLock);
    if (!scratchPtr[dev]) {
        void* p = nullptr;
        RACE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return ERR_INTERNAL;
        scratchPtr[dev] = p;
    }
    *out = scratchPtr[dev];
    return OK;
}

size_t alignUp(size_t v) { return (v + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }
}  // namespace

size_t refineStaticScratchBytes() { return kScratchBytes; }

int refineSolveStatic(DeviceFn* f, double tol, int maxOuter, hipStream_t st, int* outer, long* sweeps,
                      double* relres) {
    RefineFn* r = dynamic_cast<RefineFn*>(f);
    if (!r) return ERR_INVALID_ARG;
    const size_t vb = alignUp(r->vecBytes());
    const size_t need = kScratchAlign + 2 * vb;
    if (need > kScratchBytes) {
        std::fprintf(stderr,
                     "race_refine_noalloc: %d rows x %d right-hand sides need %zu bytes of workspace, more "
                     "than the %zu bytes of the static workspace; use race_refine\n",
                     r->n, r->nrhs, need, kScratchBytes);
        return ERR_INVALID_ARG;
    }
    void* base = nullptr;
    int e = scratchOfCurrentDevice(&base);
    if (e) return e;
    RefineWork w;
    w.norms = (double*)base;
    w.r = (unsigned char*)base + kScratchAlign;
    w.d = (unsigned char*)base + kScratchAlign + vb;
    const long before = r->totalSweeps;
    e = r->solve(w, tol, maxOuter, st, outer, relres);
    if (sweeps) *sweeps = r->totalSweeps - before;
    return e;
}

}  // namespace race
