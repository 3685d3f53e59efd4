// exec.hpp -- colour-phase launcher and device kernel functions.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../../include/race_amd.h"
#include "engine.hpp"
#include "hip_util.hpp"

namespace race {

// Dependency waves inside the leaves of a schedule.  RACE runs the rows of a
// leaf sequentially; rows of one wave have no dependency on each other, and
// running the waves in order reproduces the sequential result exactly.
struct WaveTable {
    std::vector<int> phaseLeafPtr;  // phase -> [first leaf, last leaf)
    std::vector<int> leafWavePtr;   // leaf  -> [first wave, last wave)
    std::vector<int> wavePtr;       // wave  -> [first row, last row) in rows
    std::vector<int> rows;          // row ids, wave by wave
    long critical = 0;              // sum over phases of the deepest leaf
};
// kind: RACE_K_GS (row writes its own x, reads its neighbours) or
// RACE_K_KACZ (row reads and writes x at all its columns)
WaveTable buildWaves(const Schedule& s, int kind, int n, const int* rowPtr, const int* col);

class DeviceFn {
  public:
    virtual ~DeviceFn() {}
    virtual int run(bool rev, hipStream_t st) = 0;
    virtual int waves(bool /*rev*/) const { return 0; }
    // run() only enqueues kernel launches on the stream (no host waits, no
    // memsets): it can be captured into a HIP graph
    virtual bool graphable() const { return false; }
    // kernel launches one execution issues
    virtual int launches(bool /*rev*/) const { return 1; }
    // one-launch sweeps whose dependency waits gave up (0 unless the
    // dependency table is broken; synchronises with the device)
    virtual int flowAborts(int* out) const {
        *out = 0;
        return 0;
    }
    void* x = nullptr;
    void* b = nullptr;
    int power = 1;
    int subPower = 1;        // MPK: products with A per power (setSubPower)
    int powerMax = 1 << 30;  // largest power x subPower accepted (MPK: the registered one)
    int kind = 0;
    int vecPrecision = 0;  // RACE_FP64 or RACE_FP32 vectors
    int nrow = 0;          // rows of the matrix the function was registered with
    int nrhs = 1;          // right-hand sides of its vectors
};

struct ExecContext {
    const Schedule* fwd = nullptr;
    const Schedule* rev = nullptr;
    const MPKPartition* mpk = nullptr;
    int layout = 0;           // 0: leaf bundles (SELL, LDS waves), 1: one workgroup per leaf
    int bundleRows = 128;     // rows per bundle
    int bundleLocal = 6144;   // LDS columns (local column space) per bundle
    int xcdOrder = 0;         // map consecutive bundles to the same XCD
    int bundleThreads = 256;  // workgroup size of the bundle kernels (64 ... 1024)
    int bundleTpr = 0;        // lanes per row in the bundle kernels (0: from the row length)
    int symmMode = 0;         // SymmSpMV: 0 auto, 1 colour phases, 2 SELL atomics, 3 expanded
    int sellC = 64;           // batched SELL kernels: rows per chunk (8, 16, 32, 64)
    int sellSigma = 1;        // batched SELL kernels: row-sorting window (1 = unsorted)
    int flow = 0;             // Gauss-Seidel: one launch per sweep with per-chain dependency flags
};

int makeDeviceFn(const race_kernel_desc& d, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out);

// one matrix of a batch: its colour schedules and its descriptor (host CRS
// and device vectors)
struct BatchItem {
    const Schedule* fwd = nullptr;
    const Schedule* rev = nullptr;
    race_kernel_desc d{};
};
// GS / KACZ on a batch of matrices, fused launches per phase (bundle_exec.hip)
int makeBundleBatchFn(const std::vector<BatchItem>& items, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out);
// SymmSpMV / SpMV on a batch of matrices, one launch (sell.hip)
int makeSellBatchFn(const std::vector<BatchItem>& items, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out);
// Gauss-Seidel / Kaczmarz over leaf bundles (bundle_exec.hip)
int makeBundleFn(const race_kernel_desc& d, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out);
// FP64 iterative refinement around an inner GS/KACZ function (refine.hip)
int makeRefineFn(const race_kernel_desc& d, DeviceFn* inner, int sweeps, std::unique_ptr<DeviceFn>* out);
int refineSolve(DeviceFn* f, double tol, int maxOuter, hipStream_t st, int* outer, long* sweeps,
                double* relres);
int refineResidual(DeviceFn* f, hipStream_t st, double* relres);
// refineSolve with its workspace in the library's static device array (no
// allocation; calls on one GPU share that array)
int refineSolveStatic(DeviceFn* f, double tol, int maxOuter, hipStream_t st, int* outer, long* sweeps,
                      double* relres);
size_t refineStaticScratchBytes();
// Gauss-Seidel / Kaczmarz with one workgroup per leaf (exec.hip)
int makeWaveFn(const race_kernel_desc& d, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out);
// SELL-C SpMV (one right-hand side) and level-blocked matrix powers (sell.hip)
int makeSellSpmvFn(const race_kernel_desc& d, std::unique_ptr<DeviceFn>* out);
int makeMpkFn(const race_kernel_desc& d, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out);
// r = b - A x in FP64 with the sums of r^2 and b^2 added to norms[0..1]
// (the refinement's residual, over the SELL layout: sell.hip)
class ResidualOp {
  public:
    virtual ~ResidualOp() {}
    virtual int run(const double* x, const double* b, void* r, bool rFp32, int nrhs, double* norms,
                    hipStream_t st) = 0;
};
int makeSellResidual(const race_kernel_desc& d, std::unique_ptr<ResidualOp>* out);
// SpMTV b += A^T x over the transpose formed at registration (sell.hip)
int makeSpmtvFn(const race_kernel_desc& d, std::unique_ptr<DeviceFn>* out);
// single-launch SymmSpMV: mode 2 SELL atomics, mode 3 expanded full matrix (sell.hip)
int makeSymmSellFn(const race_kernel_desc& d, int mode, std::unique_ptr<DeviceFn>* out);
// streaming device copy, bytes a multiple of 16
int deviceCopy(void* dst, const void* src, size_t bytes, hipStream_t st);

}  // namespace race
