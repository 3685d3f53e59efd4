// sell.hip -- SELL-C matrices for the row-parallel kernels: SpMV and the
// level-blocked matrix powers (MPK).
//
// A chunk is up to 64 consecutive rows (one wavefront, one row per lane);
// its entries are stored column-major, entry j of lane i at off + 64 j + i,
// so every load instruction of a wave reads one contiguous 512-byte line of
// values (256 of indices).  Rows shorter than the chunk's widest row are
// padded with value 0 on the row's own column.  Chunks never cross a "cut"
// row (the level-group boundaries for MPK), so a row range made of whole
// groups is a contiguous range of chunks.  Rows keep their order (sigma = 1:
// the level ordering is what the MPK blocking relies on).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "exec.hpp"
#include "exec_common.hpp"

namespace race {

namespace {

constexpr int kC = 64;               // rows per chunk (one wavefront)
constexpr int kWavesPerBlock = kBlock / kC;

// 16 bytes, one load: entry offset in units of 64 (every chunk stores
// 64 * width entries, so offsets are multiples of 64)
struct SellChunk {
    int row0, rows, width, off64;
    __host__ __device__ long off() const { return (long)off64 * kC; }
};

// matrix value storage (double, float, or bfloat16 bits); the arithmetic
// runs in the vector type XT: FP64, or FP32 accumulation for FP32 / BF16
template <typename VT>
VT toStorage(double v) {
    if constexpr (sizeof(VT) == 2)
        return toBf16(v);
    else
        return (VT)v;
}
template <typename XT>
__device__ __forceinline__ XT fromStorage(double v) {
    return (XT)v;
}
template <typename XT>
__device__ __forceinline__ XT fromStorage(float v) {
    return (XT)v;
}
template <typename XT>
__device__ __forceinline__ XT fromStorage(uint16_t v) {
    return (XT)__uint_as_float((unsigned)v << 16);
}

template <typename VT>
struct SellMatrix {
    int n = 0, nChunk = 0;
    int uniformW = -1;          // >= 0: every chunk has this width, chunk c at c * 64 * width
    int maxW = 0;               // the widest chunk
    bool narrow() const { return maxW <= 8; }
    long entries = 0;
    std::vector<int> cutChunk;  // chunk index of every cut
    DevBuf<SellChunk> chunks;
    DevBuf<int> col;
    DevBuf<VT> val;

    // cuts: sorted row boundaries starting with 0 and ending with n.  With
    // one range and chunk widths close to each other (padding every chunk to
    // the widest adds at most 1/8), all chunks get the widest width: the
    // kernels then compute a chunk's place instead of loading it.
    int build(const race_kernel_desc& d, const std::vector<int>& cuts, bool allowUniform = false) {
        n = d.nrow;
        const int* rp = d.rowPtr;
        std::vector<SellChunk> ch;
        cutChunk.clear();
        long total = 0;
        for (size_t k = 0; k + 1 < cuts.size(); ++k) {
            cutChunk.push_back((int)ch.size());
            for (int r0 = cuts[k]; r0 < cuts[k + 1]; r0 += kC) {
                const int rows = std::min(kC, cuts[k + 1] - r0);
                int w = 0;
                for (int i = 0; i < rows; ++i) w = std::max(w, rp[r0 + i + 1] - rp[r0 + i]);
                ch.push_back(SellChunk{r0, rows, w, (int)(total / kC)});
                total += (long)w * kC;
            }
        }
        uniformW = -1;
        if (allowUniform && cuts.size() == 2 && !ch.empty()) {
            int wmax = 0;
            for (const SellChunk& c : ch) wmax = std::max(wmax, c.width);
            const long padded = (long)wmax * kC * (long)ch.size();
            if (padded * 8 <= total * 9) {
                uniformW = wmax;
                total = 0;
                for (SellChunk& c : ch) {
                    c.width = wmax;
                    c.off64 = (int)(total / kC);
                    total += (long)wmax * kC;
                }
            }
        }
        cutChunk.push_back((int)ch.size());
        nChunk = (int)ch.size();
        entries = total;
        maxW = 0;
        for (const SellChunk& c : ch) maxW = std::max(maxW, c.width);
        if (total / kC > (long)INT32_MAX) return ERR_INVALID_ARG;  // offsets in units of 64 must fit an int
        std::vector<int> hc((size_t)total);
        std::vector<VT> hv((size_t)total);
        for (const SellChunk& c : ch)
            for (int i = 0; i < kC; ++i) {
                const int r = c.row0 + std::min(i, c.rows - 1);
                const int len = i < c.rows ? rp[r + 1] - rp[r] : 0;
                for (int j = 0; j < c.width; ++j) {
                    const long pos = c.off() + (long)j * kC + i;
                    if (j < len) {
                        hc[pos] = d.col[rp[r] + j];
                        hv[pos] = toStorage<VT>(d.val[rp[r] + j]);
                    } else {
                        hc[pos] = r;
                        hv[pos] = toStorage<VT>(0.0);
                    }
                }
            }
        int e;
        if ((e = upload(chunks, ch.data(), ch.size()))) return e;
        if ((e = upload(col, hc.data(), hc.size()))) return e;
        if ((e = upload(val, hv.data(), hv.size()))) return e;
        return OK;
    }
};

// G entries of a row from entry j0: the column ids and values are loaded
// together, then the x gathers, so each group is two dependent memory
// levels and its loads need no branch.  CLAMP (the row's last group): past
// the chunk's width a lane re-reads the chunk's last entry and multiplies
// its value by 0 (a product, not a select on the loaded value, which the
// compiler would turn into a branch around the gathers); t + (+-0) * x == t
// for finite values (t starts at +0, and a sum that starts at +0 is never
// -0).
template <int G, bool CLAMP, typename VT, typename XT>
__device__ __forceinline__ void sellGroup(const int* __restrict__ cc, const VT* __restrict__ vv, int j0, int width,
                                          const XT* __restrict__ x, XT& t) {
    int c[G];
    XT v[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int j = CLAMP ? min(j0 + u, width - 1) : j0 + u;
        c[u] = cc[j * kC];
        const XT a = fromStorage<XT>(vv[j * kC]);
        v[u] = CLAMP ? a * (j0 + u < width ? XT(1) : XT(0)) : a;
    }
#pragma unroll
    for (int u = 0; u < G; ++u) t += v[u] * x[c[u]];
}

// y[row] = sum_j a_j x[c_j] for the rows of one chunk; sequential in j, as
// the CRS row order.  NARROW (no chunk wider than 8 entries): one clamped
// group of 8, two dependent memory levels per row.  Otherwise whole groups
// of 8, then clamped groups of 4 (at most 3 entries re-read per row).  The
// two shapes are separate kernels: in one kernel the compiler merges the
// tails and gathers part of the group in a second round.  The padding lanes
// of a chunk hold value-0 entries on a real row, so every lane of the chunk
// may compute; callers guard the store.
template <bool NARROW, typename VT, typename XT>
__device__ __forceinline__ XT sellRow(const SellChunk& k, int lane, const int* __restrict__ col,
                                      const VT* __restrict__ val, const XT* __restrict__ x) {
    const int* __restrict__ cc = col + k.off() + lane;
    const VT* __restrict__ vv = val + k.off() + lane;
    XT t = XT(0);
    if (NARROW) {
        if (k.width > 0) sellGroup<8, true>(cc, vv, 0, k.width, x, t);
    } else {
        int j0 = 0;
        for (; j0 + 8 <= k.width; j0 += 8) sellGroup<8, false>(cc, vv, j0, k.width, x, t);
        for (; j0 < k.width; j0 += 4) sellGroup<4, true>(cc, vv, j0, k.width, x, t);
    }
    return t;
}

// y = A x (ACC: y += A x) for nrhs right-hand sides (vectors column-major,
// v[r + k n])
template <bool ACC, bool NARROW, typename VT, typename XT>
__global__ void __launch_bounds__(kBlock) k_sell_spmv(const SellChunk* __restrict__ chunks, int nChunk,
                                                      const int* __restrict__ col, const VT* __restrict__ val,
                                                      const XT* __restrict__ x, XT* __restrict__ y, long n,
                                                      int nrhs) {
    const int c = blockIdx.x * kWavesPerBlock + threadIdx.x / kC;
    const int lane = threadIdx.x % kC;
    if (c >= nChunk) return;
    const SellChunk k = chunks[c];
    for (int q = 0; q < nrhs; ++q) {
        const XT t = sellRow<NARROW>(k, lane, col, val, x + q * n);
        if (lane >= k.rows) continue;
        if (ACC)
            y[q * n + k.row0 + lane] += t;
        else
            y[q * n + k.row0 + lane] = t;
    }
}

// the same over uniform chunks: chunk c holds rows [64 c, 64 c + 64) at
// c * 64 * W, no chunk records to load
template <bool ACC, bool NARROW, typename VT, typename XT>
__global__ void __launch_bounds__(kBlock) k_sell_spmv_u(int W, const int* __restrict__ col,
                                                        const VT* __restrict__ val, const XT* __restrict__ x,
                                                        XT* __restrict__ y, int n, int nrhs) {
    const int c = blockIdx.x * kWavesPerBlock + threadIdx.x / kC;
    const int lane = threadIdx.x % kC;
    const int r = c * kC + lane;
    if (c * kC >= n) return;  // a whole wave past the last chunk
    const SellChunk k{0, kC, W, c * W};
    for (int q = 0; q < nrhs; ++q) {
        const XT t = sellRow<NARROW>(k, lane, col, val, x + (long)q * n);
        if (r >= n) continue;  // padding lanes of the last chunk
        if (ACC)
            y[(long)q * n + r] += t;
        else
            y[(long)q * n + r] = t;
    }
}

// SymmSpMV over the upper triangle (diagonal included) in one pass:
// b[r] += a_rr x_r + sum_{c>r} a_rc x_c and b[c] += a_rc x_r, the update of
// SYMM_SPMV_KERNEL (examples/src/helpers/kernels.cpp:403-427).  Rows of
// different leaves of a distance-2 colour never write the same b entry, but
// rows inside a leaf do, and all rows run at once here: every write to b is
// a device-scope FP64 atomic add.  Padding entries (value 0 on the row's
// own column) add nothing.
__global__ void __launch_bounds__(kBlock) k_sell_symm_atomic(const SellChunk* __restrict__ chunks, int nChunk,
                                                             const int* __restrict__ col,
                                                             const double* __restrict__ val,
                                                             const double* __restrict__ x, double* b, long n,
                                                             int nrhs) {
    const int c = blockIdx.x * kWavesPerBlock + threadIdx.x / kC;
    const int lane = threadIdx.x % kC;
    if (c >= nChunk) return;
    const SellChunk k = chunks[c];
    if (lane >= k.rows) return;
    const int r = k.row0 + lane;
    const int* __restrict__ cc = col + k.off() + lane;
    const double* __restrict__ vv = val + k.off() + lane;
    for (int q = 0; q < nrhs; ++q) {
        const double* __restrict__ xq = x + q * n;
        double* bq = b + q * n;
        const double xr = xq[r];
        double t = 0.0;
        for (int j = 0; j < k.width; ++j) {
            const int cj = cc[j * kC];
            const double v = vv[j * kC];
            if (cj == r) {
                t += v * xr;
            } else {
                t += v * xq[cj];
                atomicAdd(&bq[cj], v * xr);
            }
        }
        atomicAdd(&bq[r], t);
    }
}

// One wavefront step of the matrix powers: segment s computes power q_s on
// chunks [c0, c1) (whole level groups) from power q_s - 1.
struct MpkSeg {
    int c0, c1, q, blockStart;
};

// Up to kSegArgs segments of a step travel by value in the kernel
// arguments, so finding a block's segment costs no memory round trip; a
// step with more segments (powers above 16) reads them from the device
// table.
constexpr int kSegArgs = 16;
struct MpkSegArgs {
    MpkSeg s[kSegArgs];
};

// Power q is stored in vector slot q % nslot: nslot = k + 1 keeps every
// power, nslot = 2 keeps the last two (the only-highest variant).
template <typename VT, typename XT, bool TABLE, bool NARROW>
__global__ void __launch_bounds__(kBlock) k_sell_mpk_step(const SellChunk* __restrict__ chunks,
                                                          const int* __restrict__ col,
                                                          const VT* __restrict__ val, XT* __restrict__ x, long n,
                                                          int nslot, const MpkSeg* __restrict__ segs,
                                                          MpkSegArgs inl, int nseg) {
    // the block's segment: the last one starting at or before it (by
    // value: selects over arguments that are all loaded, no branches)
    MpkSeg sg;
    if (TABLE) {
        int s = 0;
        while (s + 1 < nseg && (int)blockIdx.x >= segs[s + 1].blockStart) ++s;
        sg = segs[s];
    } else {
        sg = inl.s[0];
#pragma unroll
        for (int q = 1; q < kSegArgs; ++q) {
            const MpkSeg o = inl.s[q];
            const bool take = q < nseg && (int)blockIdx.x >= o.blockStart;
            sg.c0 = take ? o.c0 : sg.c0;
            sg.c1 = take ? o.c1 : sg.c1;
            sg.q = take ? o.q : sg.q;
            sg.blockStart = take ? o.blockStart : sg.blockStart;
        }
    }
    const int c = sg.c0 + ((int)blockIdx.x - sg.blockStart) * kWavesPerBlock + (int)threadIdx.x / kC;
    const int lane = threadIdx.x % kC;
    if (c >= sg.c1) return;
    const SellChunk k = chunks[c];
    const XT* __restrict__ xin = x + (long)((sg.q - 1) % nslot) * n;
    const XT t = sellRow<NARROW>(k, lane, col, val, xin);
    if (lane < k.rows) x[(long)(sg.q % nslot) * n + k.row0 + lane] = t;
}

// one launch of y = A x or y += A x over a single-range SELL matrix
template <bool ACC, bool NARROW, typename VT, typename XT>
void launchSellShape(const SellMatrix<VT>& A, const void* x, void* y, int nrhs, hipStream_t st) {
    const unsigned g = (unsigned)std::max(1, (A.nChunk + kWavesPerBlock - 1) / kWavesPerBlock);
    if (A.uniformW >= 0)
        hipLaunchKernelGGL((k_sell_spmv_u<ACC, NARROW, VT, XT>), dim3(g), dim3(kBlock), 0, st, A.uniformW,
                           A.col.p, A.val.p, (const XT*)x, (XT*)y, A.n, nrhs);
    else
        hipLaunchKernelGGL((k_sell_spmv<ACC, NARROW, VT, XT>), dim3(g), dim3(kBlock), 0, st, A.chunks.p,
                           A.nChunk, A.col.p, A.val.p, (const XT*)x, (XT*)y, (long)A.n, nrhs);
}
template <bool ACC, typename VT, typename XT>
void launchSell(const SellMatrix<VT>& A, const void* x, void* y, int nrhs, hipStream_t st) {
    if (A.narrow())
        launchSellShape<ACC, true, VT, XT>(A, x, y, nrhs, st);
    else
        launchSellShape<ACC, false, VT, XT>(A, x, y, nrhs, st);
}

// b = A x
template <typename VT, typename XT>
class SellSpmvFn : public DeviceFn {
  public:
    SellMatrix<VT> A;
    int nrhs = 1;
    int init(const race_kernel_desc& d) {
        kind = d.kind;
        x = d.x;
        b = d.b;
        nrhs = d.nrhs > 0 ? d.nrhs : 1;
        return A.build(d, {0, d.nrow}, true);
    }
    int run(bool, hipStream_t st) override {
        launchSell<false, VT, XT>(A, x, b, nrhs, st);
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

// SymmSpMV b += A x from the upper triangle, in one launch:
//  mode 2: SELL-C upper triangle, FP64 atomics for every write to b (FP64 only);
//  mode 3: the full matrix (upper triangle and its transpose) built once at
//          registration, SELL-C SpMV accumulating into b (no atomics, reads
//          the strict upper part twice).
template <typename VT, typename XT>
class SymmSellFn : public DeviceFn {
  public:
    SellMatrix<VT> A;
    int mode = 2, nrhs = 1;
    int init(const race_kernel_desc& d, int mode_) {
        kind = d.kind;
        x = d.x;
        b = d.b;
        mode = mode_;
        nrhs = d.nrhs > 0 ? d.nrhs : 1;
        const int n = d.nrow;
        if (mode == 2) {
            if (!std::is_same<VT, double>::value || !std::is_same<XT, double>::value) return ERR_NOT_IMPLEMENTED;
            return A.build(d, {0, n});
        }
        // full pattern: row r holds its upper entries, then the transposed
        // entries (c, r) of rows c < r, each in increasing c
        std::vector<int> cnt(n + 1, 0);
        for (int r = 0; r < n; ++r)
            for (int j = d.rowPtr[r]; j < d.rowPtr[r + 1]; ++j) {
                ++cnt[r + 1];
                if (d.col[j] != r) ++cnt[d.col[j] + 1];
            }
        for (int r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
        std::vector<int> rp(cnt), fc(cnt[n]);
        std::vector<double> fv(cnt[n]);
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int r = 0; r < n; ++r)
            for (int j = d.rowPtr[r]; j < d.rowPtr[r + 1]; ++j) {
                fc[pos[r]] = d.col[j];
                fv[pos[r]++] = d.val[j];
            }
        for (int r = 0; r < n; ++r)
            for (int j = d.rowPtr[r]; j < d.rowPtr[r + 1]; ++j) {
                const int c = d.col[j];
                if (c == r) continue;
                fc[pos[c]] = r;
                fv[pos[c]++] = d.val[j];
            }
        race_kernel_desc f = d;
        f.rowPtr = rp.data();
        f.col = fc.data();
        f.val = fv.data();
        return A.build(f, {0, n}, true);
    }
    int run(bool, hipStream_t st) override {
        if constexpr (std::is_same<VT, double>::value && std::is_same<XT, double>::value) {
            if (mode == 2) {
                const unsigned g = (unsigned)std::max(1, (A.nChunk + kWavesPerBlock - 1) / kWavesPerBlock);
                hipLaunchKernelGGL(k_sell_symm_atomic, dim3(g), dim3(kBlock), 0, st, A.chunks.p, A.nChunk,
                                   A.col.p, A.val.p, (const double*)x, (double*)b, (long)A.n, nrhs);
                RACE_HIP(hipGetLastError());
                return OK;
            }
        }
        launchSell<true, VT, XT>(A, x, b, nrhs, st);
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

// Level-blocked matrix powers x_q = A x_{q-1}, q = 1..k, x_q at x + q n.
// Level group g at power q reads power q - 1 of groups g - 1, g, g + 1
// (BFS levels only touch neighbouring levels): the dependency rule of the
// reference's trapezoidal wavefront (src/functionManager.cpp:618-822).  Step
// L computes power q on group L - 2 (q - 1); its inputs were written by steps
// L - 3 .. L - 1, so the segments of a step are independent and run as one
// launch.  A step touches at most 2k consecutive level groups, which the
// cache model (src/matrixPower.cpp:88-140, 685-940) sized to stay in the
// Infinity Cache, so A is read from HBM about once per call.
//
// Only-highest variant (MAT_SPMV_KERNEL_only_highest
// examples/src/helpers/kernels.cpp:706-734): two vector slots, power q in
// slot q % 2.  Writing (g, q) overwrites (g, q - 2), whose readers (g - 1 .. g
// + 1 at power q - 1) ran at steps up to L - 1, and (g, q) is overwritten by
// (g, q + 2) at step L + 4 after its readers (steps up to L + 3).  Inside a
// step, a segment reads slot (q' - 1) % 2 of groups g' - 1 .. g' + 1 while
// another writes slot q % 2 of group g = g' + 2 (q' - q); the slots differ
// unless q' = q, so no segment reads what another writes.
//
// The step tables of every power 1..k are built at registration, so
// setPower(p <= k) runs the first p powers.  With FP32 / BF16 values the
// powers are FP32 vectors accumulated in FP32.
template <typename VT, typename XT>
class MpkFn : public DeviceFn {
  public:
    SellMatrix<VT> A;
    int nslot = 2;
    // for power p: segments stepSeg[p][L] .. stepSeg[p][L+1] of step L
    std::vector<std::vector<int>> stepSeg, stepBlocks;
    std::vector<MpkSeg> hostSegs;
    DevBuf<MpkSeg> segs;
    int init(const race_kernel_desc& d, const ExecContext& ctx) {
        kind = d.kind;
        x = d.x;
        b = d.b;
        if (!ctx.mpk) return ERR_INVALID_ARG;
        // powers x sub-powers: every (pow, subPow) pair of the reference's
        // power executor is one product with A, in order
        // (FuncManager::powerCallGeneral src/functionManager.cpp:618-822)
        power = d.power > 0 ? d.power : 1;
        subPower = d.sub_power > 0 ? d.sub_power : 1;
        powerMax = power * subPower;
        nslot = d.kind == RACE_K_MPK_HIGHEST ? 2 : powerMax + 1;
        const std::vector<int>& gp = ctx.mpk->groupPtr;
        if (gp.size() < 2 || gp.front() != 0 || gp.back() != d.nrow) return ERR_INVALID_ARG;
        const int G = (int)gp.size() - 1;
        // The wavefront needs every entry of group g in groups g-1..g+1.
        // BFS levels guarantee it; with mtxType "L" / "U" (row order as
        // levels) it holds only for matrices whose entries stay within
        // neighbouring groups, so anything else is refused here.
        {
            std::vector<int> grp(d.nrow);
            for (int g = 0; g < G; ++g)
                for (int r = gp[g]; r < gp[g + 1]; ++r) grp[r] = g;
            for (int r = 0; r < d.nrow; ++r)
                for (int j = d.rowPtr[r]; j < d.rowPtr[r + 1]; ++j) {
                    const int dg = grp[d.col[j]] - grp[r];
                    if (dg < -1 || dg > 1) return ERR_INCOMPATIBILITY;
                }
        }
        int e;
        if ((e = A.build(d, gp))) return e;
        std::vector<MpkSeg>& all = hostSegs;
        all.clear();
        stepSeg.assign(powerMax + 1, {});
        stepBlocks.assign(powerMax + 1, {});
        for (int p = 1; p <= powerMax; ++p) {
            stepSeg[p].assign(1, (int)all.size());
            for (int L = 0; L < G + 2 * (p - 1); ++L) {
                int blocks = 0;
                for (int q = 1; q <= p; ++q) {
                    const int g = L - 2 * (q - 1);
                    if (g < 0 || g >= G) continue;
                    const int c0 = A.cutChunk[g], c1 = A.cutChunk[g + 1];
                    if (c1 <= c0) continue;
                    all.push_back(MpkSeg{c0, c1, q, blocks});
                    blocks += (c1 - c0 + kWavesPerBlock - 1) / kWavesPerBlock;
                }
                stepSeg[p].push_back((int)all.size());
                stepBlocks[p].push_back(blocks);
            }
        }
        return upload(segs, all.data(), all.size());
    }
    bool graphable() const override { return true; }
    int launches(bool) const override {
        const int steps = power * subPower;
        if (steps < 1 || steps > powerMax) return 0;
        int k = 0;
        for (int nb : stepBlocks[steps]) k += nb > 0;
        return k;
    }
    int run(bool, hipStream_t st) override {
        const int steps = power * subPower;
        if (steps < 1 || steps > powerMax) return ERR_INVALID_ARG;
        const std::vector<int>& sg = stepSeg[steps];
        const std::vector<int>& nb = stepBlocks[steps];
        for (int L = 0; L + 1 < (int)sg.size(); ++L) {
            if (nb[L] == 0) continue;
            const int ns = sg[L + 1] - sg[L];
            MpkSegArgs inl{};
            if (ns <= kSegArgs) std::copy(hostSegs.begin() + sg[L], hostSegs.begin() + sg[L + 1], inl.s);
            const dim3 grid(nb[L]), block(kBlock);
            const MpkSeg* tab = segs.p + sg[L];
            if (ns <= kSegArgs && A.narrow())
                hipLaunchKernelGGL((k_sell_mpk_step<VT, XT, false, true>), grid, block, 0, st, A.chunks.p, A.col.p,
                                   A.val.p, (XT*)x, (long)A.n, nslot, tab, inl, ns);
            else if (ns <= kSegArgs)
                hipLaunchKernelGGL((k_sell_mpk_step<VT, XT, false, false>), grid, block, 0, st, A.chunks.p,
                                   A.col.p, A.val.p, (XT*)x, (long)A.n, nslot, tab, inl, ns);
            else if (A.narrow())
                hipLaunchKernelGGL((k_sell_mpk_step<VT, XT, true, true>), grid, block, 0, st, A.chunks.p, A.col.p,
                                   A.val.p, (XT*)x, (long)A.n, nslot, tab, inl, ns);
            else
                hipLaunchKernelGGL((k_sell_mpk_step<VT, XT, true, false>), grid, block, 0, st, A.chunks.p,
                                   A.col.p, A.val.p, (XT*)x, (long)A.n, nslot, tab, inl, ns);
        }
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

// SpMTV b += A^T x for nrhs right-hand sides (SPMTV_KERNEL
// examples/src/helpers/kernels.cpp:106-124, which scatters
// b[c_j] += a_j x[r] row by row under a distance-2 colouring).  The
// transpose is formed once at registration, so a call is one gather SpMV
// over the SELL-C rows of A^T accumulating into b: no atomics, no colour
// phases, every b entry written once.  Row c of A^T holds its entries in
// increasing r, the order a sequential sweep adds them in.
template <typename VT, typename XT>
class SpmtvFn : public DeviceFn {
  public:
    SellMatrix<VT> A;
    int nrhs = 1;
    int init(const race_kernel_desc& d) {
        kind = d.kind;
        x = d.x;
        b = d.b;
        nrhs = d.nrhs > 0 ? d.nrhs : 1;
        const int n = d.nrow;
        std::vector<int> rp(n + 1, 0);
        for (int j = 0; j < d.rowPtr[n]; ++j) ++rp[d.col[j] + 1];
        for (int r = 0; r < n; ++r) rp[r + 1] += rp[r];
        std::vector<int> tc(rp[n]), pos(rp.begin(), rp.end() - 1);
        std::vector<double> tv(rp[n]);
        for (int r = 0; r < n; ++r)
            for (int j = d.rowPtr[r]; j < d.rowPtr[r + 1]; ++j) {
                const int c = d.col[j];
                tc[pos[c]] = r;
                tv[pos[c]++] = d.val[j];
            }
        race_kernel_desc t = d;
        t.rowPtr = rp.data();
        t.col = tc.data();
        t.val = tv.data();
        return A.build(t, {0, n}, true);
    }
    int run(bool, hipStream_t st) override {
        launchSell<true, VT, XT>(A, x, b, nrhs, st);
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

// r = b - A x (FP64 product, r stored as RT) and the sums of r^2 and b^2
// for the refinement; grid-stride over chunks so few workgroups add to the
// two norm accumulators
template <typename RT, bool NARROW>
__global__ void __launch_bounds__(kBlock) k_sell_residual(const SellChunk* __restrict__ chunks, int nChunk,
                                                          int W, const int* __restrict__ col,
                                                          const double* __restrict__ val,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ b, RT* __restrict__ r,
                                                          long n, int nrhs, double* __restrict__ norms) {
    __shared__ double sr[kWavesPerBlock], sb[kWavesPerBlock];
    const int lane = threadIdx.x % kC, wv = threadIdx.x / kC;
    double rr = 0.0, bb = 0.0;
    for (int c = blockIdx.x * kWavesPerBlock + wv; c < nChunk; c += gridDim.x * kWavesPerBlock) {
        const SellChunk k =
            W >= 0 ? SellChunk{c * kC, (int)std::min<long>(kC, n - (long)c * kC), W, c * W} : chunks[c];
        const bool live = lane < k.rows;  // padding lanes compute, store nothing
        const long row = k.row0 + (live ? lane : 0);
        for (int q = 0; q < nrhs; ++q) {
            const double t = sellRow<NARROW>(k, lane, col, val, x + q * n);
            const double bv = b[q * n + row];
            if (!live) continue;
            const double rv = bv - t;
            r[q * n + row] = (RT)rv;
            rr += rv * rv;
            bb += bv * bv;
        }
    }
    rr = groupSum<kC>(rr);
    bb = groupSum<kC>(bb);
    if (lane == 0) {
        sr[wv] = rr;
        sb[wv] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, e = 0.0;
        for (int w = 0; w < kWavesPerBlock; ++w) {
            a += sr[w];
            e += sb[w];
        }
        atomicAdd(&norms[0], a);
        atomicAdd(&norms[1], e);
    }
}

class SellResidual : public ResidualOp {
  public:
    SellMatrix<double> A;
    int init(const race_kernel_desc& d) { return A.build(d, {0, d.nrow}, true); }
    template <typename RT, bool NARROW>
    void launch(const double* x, const double* b, void* r, int nrhs, double* norms, hipStream_t st) {
        const int blocks = std::max(1, std::min(4096, (A.nChunk + kWavesPerBlock - 1) / kWavesPerBlock));
        hipLaunchKernelGGL((k_sell_residual<RT, NARROW>), dim3(blocks), dim3(kBlock), 0, st, A.chunks.p, A.nChunk,
                           A.uniformW, A.col.p, A.val.p, x, b, (RT*)r, (long)A.n, nrhs, norms);
    }
    int run(const double* x, const double* b, void* r, bool rFp32, int nrhs, double* norms,
            hipStream_t st) override {
        if (rFp32 && A.narrow())
            launch<float, true>(x, b, r, nrhs, norms, st);
        else if (rFp32)
            launch<float, false>(x, b, r, nrhs, norms, st);
        else if (A.narrow())
            launch<double, true>(x, b, r, nrhs, norms, st);
        else
            launch<double, false>(x, b, r, nrhs, norms, st);
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

// ---------------------------------------------------------------------------
// Generalised SELL-C-sigma for batches and the layout tuner: chunk height
// C in {8, 16, 32, 64} (a wavefront holds 64 / C chunks), rows sorted by
// length (longest first) inside windows of sigma rows so that a chunk holds
// rows of similar length, and every chunk tagged with its matrix, whose
// vectors it uses.  Entry j of slot i of a chunk at off + j C + i.
struct SellGChunk {
    int rows, width, slot0, mat;
    long off;
};

template <int C, bool PERM, typename XT, bool ACC>
__global__ void __launch_bounds__(kBlock) k_sellg_spmv(const SellGChunk* __restrict__ chunks, int nChunk,
                                                       const int* __restrict__ rowOf, const int* __restrict__ col,
                                                       const XT* __restrict__ val, const XT* const* __restrict__ mx,
                                                       XT* const* __restrict__ my, const int* __restrict__ mn,
                                                       int nrhs) {
    const int lane = threadIdx.x % 64;
    const int c = (int)((blockIdx.x * (long)blockDim.x + threadIdx.x) / 64) * (64 / C) + lane / C;
    const int i = lane % C;
    if (c >= nChunk) return;
    const SellGChunk k = chunks[c];
    if (i >= k.rows) return;
    const int row = PERM ? rowOf[k.slot0 + i] : k.slot0 + i;
    const XT* __restrict__ x = mx[k.mat];
    XT* __restrict__ y = my[k.mat];
    const long n = mn[k.mat];
    const int* __restrict__ cc = col + k.off + i;
    const XT* __restrict__ vv = val + k.off + i;
    for (int q = 0; q < nrhs; ++q) {
        const XT* __restrict__ xq = x + q * n;
        XT t = XT(0);
        int j = 0;
        for (; j + 2 <= k.width; j += 2) {
            const int c0 = cc[j * C], c1 = cc[(j + 1) * C];
            const XT v0 = vv[j * C], v1 = vv[(j + 1) * C];
            t += v0 * xq[c0];
            t += v1 * xq[c1];
        }
        if (j < k.width) t += vv[j * C] * xq[cc[j * C]];
        if (ACC)
            y[q * n + row] += t;
        else
            y[q * n + row] = t;
    }
}

// full CRS of a symmetric matrix from its upper triangle (diagonal first):
// row r holds its upper entries, then (c, r) of the rows c < r
static void expandUpper(const race_kernel_desc& d, std::vector<int>& rp, std::vector<int>& fc,
                        std::vector<double>& fv) {
    const int n = d.nrow;
    std::vector<int> cnt(n + 1, 0);
    for (int r = 0; r < n; ++r)
        for (int j = d.rowPtr[r]; j < d.rowPtr[r + 1]; ++j) {
            ++cnt[r + 1];
            if (d.col[j] != r) ++cnt[d.col[j] + 1];
        }
    for (int r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
    rp = cnt;
    fc.assign(cnt[n], 0);
    fv.assign(cnt[n], 0.0);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int r = 0; r < n; ++r)
        for (int j = d.rowPtr[r]; j < d.rowPtr[r + 1]; ++j) {
            fc[pos[r]] = d.col[j];
            fv[pos[r]++] = d.val[j];
        }
    for (int r = 0; r < n; ++r)
        for (int j = d.rowPtr[r]; j < d.rowPtr[r + 1]; ++j) {
            const int c = d.col[j];
            if (c == r) continue;
            fc[pos[c]] = r;
            fv[pos[c]++] = d.val[j];
        }
}

// SpMV (b = A x) or SymmSpMV (b += A x, expanded from the upper triangle)
// on a batch of matrices in one launch.
template <typename XT>
class SellBatchFn : public DeviceFn {
  public:
    int C = 64, sigma = 1, nrhs = 1, nChunk = 0;
    bool perm = false;
    DevBuf<SellGChunk> chunks;
    DevBuf<int> rowOf, col, mn;
    DevBuf<XT> val;
    DevBuf<const XT*> mx;
    DevBuf<XT*> my;

    int init(const std::vector<BatchItem>& items, const ExecContext& ctx) {
        kind = items[0].d.kind;
        nrhs = items[0].d.nrhs > 0 ? items[0].d.nrhs : 1;
        C = ctx.sellC;
        sigma = std::max(1, ctx.sellSigma);
        perm = sigma > 1;
        if (C != 8 && C != 16 && C != 32 && C != 64) return ERR_INVALID_ARG;
        std::vector<SellGChunk> hc;
        std::vector<int> hrow, hcol, hn;
        std::vector<XT> hval;
        std::vector<const XT*> hx;
        std::vector<XT*> hy;
        std::vector<int> rp, fc, order;
        std::vector<double> fv;
        for (size_t m = 0; m < items.size(); ++m) {
            const race_kernel_desc& d = items[m].d;
            const int n = d.nrow;
            const int* R = d.rowPtr;
            const int* Cc = d.col;
            const double* V = d.val;
            if (kind == RACE_K_SYMMSPMV) {
                expandUpper(d, rp, fc, fv);
                R = rp.data();
                Cc = fc.data();
                V = fv.data();
            }
            order.resize(n);
            for (int r = 0; r < n; ++r) order[r] = r;
            if (perm)
                for (int w0 = 0; w0 < n; w0 += sigma) {
                    const int w1 = std::min(n, w0 + sigma);
                    std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int a, int b) {
                        return R[a + 1] - R[a] > R[b + 1] - R[b];
                    });
                }
            for (int r0 = 0; r0 < n; r0 += C) {
                const int rows = std::min(C, n - r0);
                int w = 0;
                for (int i = 0; i < rows; ++i) w = std::max(w, R[order[r0 + i] + 1] - R[order[r0 + i]]);
                const long off = (long)hcol.size();
                const int slot0 = perm ? (int)hrow.size() : r0;
                hc.push_back(SellGChunk{rows, w, slot0, (int)m, off});
                hcol.resize(off + (long)w * C);
                hval.resize(off + (long)w * C);
                for (int i = 0; i < C; ++i) {
                    const int r = order[r0 + std::min(i, rows - 1)];
                    const int len = i < rows ? R[r + 1] - R[r] : 0;
                    for (int j = 0; j < w; ++j) {
                        const long pos = off + (long)j * C + i;
                        hcol[pos] = j < len ? Cc[R[r] + j] : r;
                        hval[pos] = j < len ? (XT)V[R[r] + j] : XT(0);
                    }
                    if (perm && i < rows) hrow.push_back(r);
                }
            }
            hx.push_back((const XT*)d.x);
            hy.push_back((XT*)d.b);
            hn.push_back(n);
        }
        nChunk = (int)hc.size();
        int e;
        if ((e = upload(chunks, hc.data(), hc.size()))) return e;
        if ((e = upload(rowOf, hrow.data(), hrow.size()))) return e;
        if ((e = upload(col, hcol.data(), hcol.size()))) return e;
        if ((e = upload(val, hval.data(), hval.size()))) return e;
        if ((e = upload(mx, hx.data(), hx.size()))) return e;
        if ((e = upload(my, hy.data(), hy.size()))) return e;
        if ((e = upload(mn, hn.data(), hn.size()))) return e;
        return OK;
    }
    template <int CC, bool P>
    void launch(hipStream_t st) {
        const long waves = ((long)nChunk + (64 / CC) - 1) / (64 / CC);
        const unsigned g = (unsigned)std::max(1l, (waves + kWavesPerBlock - 1) / kWavesPerBlock);
        if (kind == RACE_K_SYMMSPMV)
            hipLaunchKernelGGL((k_sellg_spmv<CC, P, XT, true>), dim3(g), dim3(kBlock), 0, st, chunks.p, nChunk,
                               rowOf.p, col.p, val.p, mx.p, my.p, mn.p, nrhs);
        else
            hipLaunchKernelGGL((k_sellg_spmv<CC, P, XT, false>), dim3(g), dim3(kBlock), 0, st, chunks.p, nChunk,
                               rowOf.p, col.p, val.p, mx.p, my.p, mn.p, nrhs);
    }
    int run(bool, hipStream_t st) override {
        switch (C * 2 + (perm ? 1 : 0)) {
            case 16: launch<8, false>(st); break;
            case 17: launch<8, true>(st); break;
            case 32: launch<16, false>(st); break;
            case 33: launch<16, true>(st); break;
            case 64: launch<32, false>(st); break;
            case 65: launch<32, true>(st); break;
            case 128: launch<64, false>(st); break;
            default: launch<64, true>(st); break;
        }
        RACE_HIP(hipGetLastError());
        return OK;
    }
};

}  // namespace

int makeSellBatchFn(const std::vector<BatchItem>& items, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out) {
    if (items.empty()) return ERR_INVALID_ARG;
    const race_kernel_desc& d0 = items[0].d;
    for (const BatchItem& it : items)
        if (it.d.kind != d0.kind || it.d.val_precision != d0.val_precision ||
            it.d.vec_precision != d0.vec_precision || it.d.nrhs != d0.nrhs)
            return ERR_INVALID_ARG;
    if (d0.kind != RACE_K_SYMMSPMV && d0.kind != RACE_K_SPMV) return ERR_INVALID_ARG;
    if (d0.val_precision != d0.vec_precision) return ERR_NOT_IMPLEMENTED;
    int e;
    if (d0.vec_precision == RACE_FP64) {
        auto f = std::make_unique<SellBatchFn<double>>();
        e = f->init(items, ctx);
        *out = std::move(f);
    } else if (d0.vec_precision == RACE_FP32) {
        auto f = std::make_unique<SellBatchFn<float>>();
        e = f->init(items, ctx);
        *out = std::move(f);
    } else {
        return ERR_NOT_IMPLEMENTED;
    }
    return e;
}

// the storage / arithmetic pairs of the single-matrix SELL functions:
// FP64 / FP64, FP32 / FP32, BF16 / FP32
template <template <typename, typename> class F, typename Init>
int makeTyped(const race_kernel_desc& d, std::unique_ptr<DeviceFn>* out, Init&& init) {
    int e;
    if (d.val_precision == RACE_FP64 && d.vec_precision == RACE_FP64) {
        auto f = std::make_unique<F<double, double>>();
        e = init(*f);
        *out = std::move(f);
    } else if (d.val_precision == RACE_FP32 && d.vec_precision == RACE_FP32) {
        auto f = std::make_unique<F<float, float>>();
        e = init(*f);
        *out = std::move(f);
    } else if (d.val_precision == RACE_BF16 && d.vec_precision == RACE_FP32) {
        auto f = std::make_unique<F<uint16_t, float>>();
        e = init(*f);
        *out = std::move(f);
    } else {
        return ERR_NOT_IMPLEMENTED;
    }
    return e;
}

int makeSellResidual(const race_kernel_desc& d, std::unique_ptr<ResidualOp>* out) {
    auto f = std::make_unique<SellResidual>();
    int e = f->init(d);
    *out = std::move(f);
    return e;
}

int makeSellSpmvFn(const race_kernel_desc& d, std::unique_ptr<DeviceFn>* out) {
    return makeTyped<SellSpmvFn>(d, out, [&](auto& f) { return f.init(d); });
}

int makeSymmSellFn(const race_kernel_desc& d, int mode, std::unique_ptr<DeviceFn>* out) {
    return makeTyped<SymmSellFn>(d, out, [&](auto& f) { return f.init(d, mode); });
}

int makeSpmtvFn(const race_kernel_desc& d, std::unique_ptr<DeviceFn>* out) {
    return makeTyped<SpmtvFn>(d, out, [&](auto& f) { return f.init(d); });
}

int makeMpkFn(const race_kernel_desc& d, const ExecContext& ctx, std::unique_ptr<DeviceFn>* out) {
    return makeTyped<MpkFn>(d, out, [&](auto& f) { return f.init(d, ctx); });
}

}  // namespace race
