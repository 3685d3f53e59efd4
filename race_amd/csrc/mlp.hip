// mlp.hip -- the layout tuner's model: a 16-64-64-12 multilayer perceptron
// (ReLU hidden layers, linear output) that maps 16 matrix features to the
// predicted relative log-time of the 12 candidate layouts.
//
// Training is FP32 on the device: one thread per output element for the
// dense layers and their gradients (dW = X^T dY reduces over the batch in
// a fixed order, so gradients are deterministic), masked mean-squared
// error, Adam.  Inference has an FP32 path and a BF16 path on the matrix
// cores: one wavefront computes 16 samples with v_mfma_f32_16x16x16_bf16
// tiles (FP32 accumulation), activations staged through LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/race_amd.h"
#include "engine.hpp"
#include "exec_common.hpp"
#include "hip_util.hpp"

namespace race {

namespace {

constexpr int kIn = 16, kH = 64, kOut = 12, kOutPad = 16;
// parameter block: W1 [kIn][kH], c1 [kH], W2 [kH][kH], c2 [kH], W3 [kH][kOut], c3 [kOut]
constexpr int oW1 = 0, oC1 = oW1 + kIn * kH, oW2 = oC1 + kH, oC2 = oW2 + kH * kH, oW3 = oC2 + kH,
              oC3 = oW3 + kH * kOut, kParams = oC3 + kOut;

// Y[b][o] = act(c[o] + sum_i X[b][i] W[i][o])
__global__ void k_dense(const float* __restrict__ X, const float* __restrict__ W, const float* __restrict__ c,
                        float* __restrict__ Y, int B, int nin, int nout, int relu) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * nout) return;
    const int b = idx / nout, o = idx % nout;
    float acc = c[o];
    for (int i = 0; i < nin; ++i) acc += X[b * nin + i] * W[i * nout + o];
    Y[idx] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dY = 2 mask (Y - T) / sum(mask); loss += mask (Y - T)^2 / sum(mask)
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ mask,
                      float* __restrict__ dY, float* __restrict__ loss, int n, float inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = (Y[i] - T[i]) * mask[i];
    dY[i] = 2.0f * d * inv;
    atomicAdd(loss, d * d * inv);
}

// dX[b][i] = (sum_o dY[b][o] W[i][o]) * (A[b][i] > 0), A the ReLU output
// that fed this layer
__global__ void k_dense_bwd_x(const float* __restrict__ dY, const float* __restrict__ W,
                              const float* __restrict__ A, float* __restrict__ dX, int B, int nin, int nout) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * nin) return;
    const int b = idx / nin, i = idx % nin;
    float acc = 0.0f;
    for (int o = 0; o < nout; ++o) acc += dY[b * nout + o] * W[i * nout + o];
    dX[idx] = A[idx] > 0.0f ? acc : 0.0f;
}

// dW[i][o] = sum_b X[b][i] dY[b][o] (i < nin); dc[o] = sum_b dY[b][o]
__global__ void k_dense_bwd_w(const float* __restrict__ X, const float* __restrict__ dY, float* __restrict__ dW,
                              float* __restrict__ dc, int B, int nin, int nout) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (nin + 1) * nout) return;
    const int i = idx / nout, o = idx % nout;
    float acc = 0.0f;
    if (i < nin)
        for (int b = 0; b < B; ++b) acc += X[b * nin + i] * dY[b * nout + o];
    else
        for (int b = 0; b < B; ++b) acc += dY[b * nout + o];
    if (i < nin)
        dW[i * nout + o] = acc;
    else
        dc[o] = acc;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ---- BF16 inference on the matrix cores ------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short toBf16Bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (short)(u >> 16);
}

// One wavefront, 16 samples.  v_mfma_f32_16x16x16_bf16 operand layout:
// A (16 x 16, rows x k): lane l holds A[l % 16][4 (l / 16) + j], j = 0..3;
// B (16 x 16, k x cols): lane l holds B[4 (l / 16) + j][l % 16];
// D (16 x 16): lane l holds D[4 (l / 16) + j][l % 16].
// Weights are BF16, row-major [in][out], W3 padded to 16 output columns.
__global__ void __launch_bounds__(64) k_mlp_bf16(const short* __restrict__ X, const short* __restrict__ W1,
                                                 const float* __restrict__ c1, const short* __restrict__ W2,
                                                 const float* __restrict__ c2, const short* __restrict__ W3,
                                                 const float* __restrict__ c3, float* __restrict__ Y, int B) {
    __shared__ short h[16][kH];
    const int l = threadIdx.x, r = l % 16, g = l / 16;
    const int b0 = blockIdx.x * 16;
    // layer 1: [16 x 16] x [16 x 64]
    bf16x4 a;
    for (int j = 0; j < 4; ++j) a[j] = (b0 + r < B) ? X[(b0 + r) * kIn + 4 * g + j] : (short)0;
    for (int t = 0; t < kH / 16; ++t) {
        bf16x4 w;
        for (int j = 0; j < 4; ++j) w[j] = W1[(4 * g + j) * kH + 16 * t + r];
        floatx4 d = {0.0f, 0.0f, 0.0f, 0.0f};
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, d, 0, 0, 0);
        for (int j = 0; j < 4; ++j) h[4 * g + j][16 * t + r] = toBf16Bits(fmaxf(d[j] + c1[16 * t + r], 0.0f));
    }
    __syncthreads();
    // layer 2: [16 x 64] x [64 x 64]
    floatx4 acc[kH / 16];
    for (int t = 0; t < kH / 16; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < kH / 16; ++ks) {
        for (int j = 0; j < 4; ++j) a[j] = h[r][16 * ks + 4 * g + j];
        for (int t = 0; t < kH / 16; ++t) {
            bf16x4 w;
            for (int j = 0; j < 4; ++j) w[j] = W2[(16 * ks + 4 * g + j) * kH + 16 * t + r];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, acc[t], 0, 0, 0);
        }
    }
    __syncthreads();  // every lane has read layer 1 from LDS
    for (int t = 0; t < kH / 16; ++t)
        for (int j = 0; j < 4; ++j) h[4 * g + j][16 * t + r] = toBf16Bits(fmaxf(acc[t][j] + c2[16 * t + r], 0.0f));
    __syncthreads();
    // layer 3: [16 x 64] x [64 x 16] (12 columns used)
    floatx4 y = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < kH / 16; ++ks) {
        for (int j = 0; j < 4; ++j) a[j] = h[r][16 * ks + 4 * g + j];
        bf16x4 w;
        for (int j = 0; j < 4; ++j) w[j] = W3[(16 * ks + 4 * g + j) * kOutPad + r];
        y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, y, 0, 0, 0);
    }
    if (r < kOut)
        for (int j = 0; j < 4; ++j)
            if (b0 + 4 * g + j < B) Y[(b0 + 4 * g + j) * kOut + r] = y[j] + c3[r];
}

inline unsigned blocks(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

short hostBf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

}  // namespace

struct MlpModel {
    DevBuf<float> p, g, m, v;  // parameters, gradients, Adam moments
    DevBuf<float> x, t, mask, h1, h2, y, dy, d2, d1, loss;
    DevBuf<short> xb, w1b, w2b, w3b;
    int cap = 0;
    long step = 0;
    hipStream_t st = nullptr;

    int reserve(int B) {
        if (B <= cap) return OK;
        RACE_HIP(x.reserve((size_t)B * kIn));
        RACE_HIP(t.reserve((size_t)B * kOut));
        RACE_HIP(mask.reserve((size_t)B * kOut));
        RACE_HIP(h1.reserve((size_t)B * kH));
        RACE_HIP(h2.reserve((size_t)B * kH));
        RACE_HIP(y.reserve((size_t)B * kOut));
        RACE_HIP(dy.reserve((size_t)B * kOut));
        RACE_HIP(d2.reserve((size_t)B * kH));
        RACE_HIP(d1.reserve((size_t)B * kH));
        RACE_HIP(xb.reserve((size_t)B * kIn));
        cap = B;
        return OK;
    }
    void forward32(int B) {
        hipLaunchKernelGGL(k_dense, dim3(blocks((long)B * kH)), dim3(256), 0, st, x.p, p.p + oW1, p.p + oC1, h1.p, B,
                           kIn, kH, 1);
        hipLaunchKernelGGL(k_dense, dim3(blocks((long)B * kH)), dim3(256), 0, st, h1.p, p.p + oW2, p.p + oC2, h2.p, B,
                           kH, kH, 1);
        hipLaunchKernelGGL(k_dense, dim3(blocks((long)B * kOut)), dim3(256), 0, st, h2.p, p.p + oW3, p.p + oC3, y.p,
                           B, kH, kOut, 0);
    }
    // gradients of the masked MSE into g (forward must have run)
    void backward(int B, float inv) {
        const int n = B * kOut;
        hipLaunchKernelGGL(k_mse, dim3(blocks(n)), dim3(256), 0, st, y.p, t.p, mask.p, dy.p, loss.p, n, inv);
        hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((kH + 1) * kOut)), dim3(256), 0, st, h2.p, dy.p, g.p + oW3,
                           g.p + oC3, B, kH, kOut);
        hipLaunchKernelGGL(k_dense_bwd_x, dim3(blocks((long)B * kH)), dim3(256), 0, st, dy.p, p.p + oW3, h2.p, d2.p,
                           B, kH, kOut);
        hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((kH + 1) * kH)), dim3(256), 0, st, h1.p, d2.p, g.p + oW2,
                           g.p + oC2, B, kH, kH);
        hipLaunchKernelGGL(k_dense_bwd_x, dim3(blocks((long)B * kH)), dim3(256), 0, st, d2.p, p.p + oW2, h1.p, d1.p,
                           B, kH, kH);
        hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((kIn + 1) * kH)), dim3(256), 0, st, x.p, d1.p, g.p + oW1,
                           g.p + oC1, B, kIn, kH);
    }
};

}  // namespace race

using race::MlpModel;

struct race_mlp_s {
    MlpModel m;
};

namespace {

int uploadBatch(MlpModel& M, const float* x, const float* t, const float* mask, int B) {
    RACE_HIP(hipMemcpyAsync(M.x.p, x, sizeof(float) * B * race::kIn, hipMemcpyHostToDevice, M.st));
    if (t) RACE_HIP(hipMemcpyAsync(M.t.p, t, sizeof(float) * B * race::kOut, hipMemcpyHostToDevice, M.st));
    if (mask) RACE_HIP(hipMemcpyAsync(M.mask.p, mask, sizeof(float) * B * race::kOut, hipMemcpyHostToDevice, M.st));
    return race::OK;
}

float maskSum(const float* mask, int n) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += mask[i];
    return (float)s;
}

}  // namespace

extern "C" {

int race_mlp_dims(int* n_in, int* n_hidden, int* n_out, int* n_params) {
    if (n_in) *n_in = race::kIn;
    if (n_hidden) *n_hidden = race::kH;
    if (n_out) *n_out = race::kOut;
    if (n_params) *n_params = race::kParams;
    return race::OK;
}

int race_mlp_create(race_mlp* out, unsigned seed) {
    if (!out) return race::ERR_INVALID_ARG;
    *out = nullptr;
    auto h = new race_mlp_s();
    MlpModel& M = h->m;
    std::vector<float> p(race::kParams, 0.0f);
    std::mt19937 gen(seed);
    auto he = [&](int off, int nin, int nout) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / nin));
        for (int i = 0; i < nin * nout; ++i) p[off + i] = nd(gen);
    };
    he(race::oW1, race::kIn, race::kH);
    he(race::oW2, race::kH, race::kH);
    he(race::oW3, race::kH, race::kOut);
    int e = race::OK;
    if (hipStreamCreateWithFlags(&M.st, hipStreamNonBlocking) != hipSuccess) e = race::ERR_INTERNAL;
    if (!e && (M.p.reserve(race::kParams) || M.g.reserve(race::kParams) || M.m.reserve(race::kParams) ||
               M.v.reserve(race::kParams) || M.loss.reserve(1)))
        e = race::ERR_INTERNAL;
    if (!e && (hipMemcpy(M.p.p, p.data(), sizeof(float) * race::kParams, hipMemcpyHostToDevice) ||
               hipMemset(M.m.p, 0, sizeof(float) * race::kParams) ||
               hipMemset(M.v.p, 0, sizeof(float) * race::kParams) ||
               hipMemset(M.g.p, 0, sizeof(float) * race::kParams)))
        e = race::ERR_INTERNAL;
    if (e) {
        if (M.st) (void)hipStreamDestroy(M.st);
        delete h;
        return e;
    }
    *out = h;
    return race::OK;
}

int race_mlp_set_params(race_mlp h, const float* params, int count) {
    if (!h || !params || count != race::kParams) return race::ERR_INVALID_ARG;
    return hipMemcpy(h->m.p.p, params, sizeof(float) * count, hipMemcpyHostToDevice) == hipSuccess
               ? race::OK
               : race::ERR_INTERNAL;
}

int race_mlp_get_params(race_mlp h, float* params, int count) {
    if (!h || !params || count != race::kParams) return race::ERR_INVALID_ARG;
    return hipMemcpy(params, h->m.p.p, sizeof(float) * count, hipMemcpyDeviceToHost) == hipSuccess
               ? race::OK
               : race::ERR_INTERNAL;
}

int race_mlp_forward(race_mlp h, const float* x, int batch, float* y, int precision) {
    if (!h || !x || !y || batch < 1) return race::ERR_INVALID_ARG;
    if (precision != RACE_FP32 && precision != RACE_BF16) return race::ERR_INVALID_ARG;
    MlpModel& M = h->m;
    int e;
    if ((e = M.reserve(batch))) return e;
    if ((e = uploadBatch(M, x, nullptr, nullptr, batch))) return e;
    if (precision == RACE_FP32) {
        M.forward32(batch);
    } else {
        // weights to BF16 (W3 padded to 16 columns), inputs to BF16
        std::vector<float> p(race::kParams);
        RACE_HIP(hipMemcpy(p.data(), M.p.p, sizeof(float) * race::kParams, hipMemcpyDeviceToHost));
        std::vector<short> w1(race::kIn * race::kH), w2(race::kH * race::kH), w3(race::kH * race::kOutPad, 0);
        for (int i = 0; i < race::kIn * race::kH; ++i) w1[i] = race::hostBf16(p[race::oW1 + i]);
        for (int i = 0; i < race::kH * race::kH; ++i) w2[i] = race::hostBf16(p[race::oW2 + i]);
        for (int i = 0; i < race::kH; ++i)
            for (int o = 0; o < race::kOut; ++o)
                w3[i * race::kOutPad + o] = race::hostBf16(p[race::oW3 + i * race::kOut + o]);
        std::vector<short> xb((size_t)batch * race::kIn);
        for (size_t i = 0; i < xb.size(); ++i) xb[i] = race::hostBf16(x[i]);
        if ((e = race::upload(M.w1b, w1.data(), w1.size())) || (e = race::upload(M.w2b, w2.data(), w2.size())) ||
            (e = race::upload(M.w3b, w3.data(), w3.size())))
            return e;
        RACE_HIP(hipMemcpyAsync(M.xb.p, xb.data(), sizeof(short) * xb.size(), hipMemcpyHostToDevice, M.st));
        hipLaunchKernelGGL(race::k_mlp_bf16, dim3((batch + 15) / 16), dim3(64), 0, M.st, M.xb.p, M.w1b.p,
                           M.p.p + race::oC1, M.w2b.p, M.p.p + race::oC2, M.w3b.p, M.p.p + race::oC3, M.y.p, batch);
    }
    RACE_HIP(hipGetLastError());
    RACE_HIP(hipMemcpyAsync(y, M.y.p, sizeof(float) * batch * race::kOut, hipMemcpyDeviceToHost, M.st));
    RACE_HIP(hipStreamSynchronize(M.st));
    return race::OK;
}

int race_mlp_gradients(race_mlp h, const float* x, const float* t, const float* mask, int batch, float* grads,
                       float* loss) {
    if (!h || !x || !t || !mask || !grads || batch < 1) return race::ERR_INVALID_ARG;
    MlpModel& M = h->m;
    int e;
    if ((e = M.reserve(batch))) return e;
    if ((e = uploadBatch(M, x, t, mask, batch))) return e;
    const float ms = maskSum(mask, batch * race::kOut);
    if (ms <= 0) return race::ERR_INVALID_ARG;
    RACE_HIP(hipMemsetAsync(M.loss.p, 0, sizeof(float), M.st));
    M.forward32(batch);
    M.backward(batch, 1.0f / ms);
    RACE_HIP(hipGetLastError());
    RACE_HIP(hipMemcpyAsync(grads, M.g.p, sizeof(float) * race::kParams, hipMemcpyDeviceToHost, M.st));
    float l = 0;
    RACE_HIP(hipMemcpyAsync(&l, M.loss.p, sizeof(float), hipMemcpyDeviceToHost, M.st));
    RACE_HIP(hipStreamSynchronize(M.st));
    if (loss) *loss = l;
    return race::OK;
}

int race_mlp_train(race_mlp h, const float* x, const float* t, const float* mask, int batch, int epochs, float lr,
                   float* loss) {
    if (!h || !x || !t || !mask || batch < 1 || epochs < 1 || !(lr > 0)) return race::ERR_INVALID_ARG;
    MlpModel& M = h->m;
    int e;
    if ((e = M.reserve(batch))) return e;
    if ((e = uploadBatch(M, x, t, mask, batch))) return e;
    const float ms = maskSum(mask, batch * race::kOut);
    if (ms <= 0) return race::ERR_INVALID_ARG;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        RACE_HIP(hipMemsetAsync(M.loss.p, 0, sizeof(float), M.st));
        M.forward32(batch);
        M.backward(batch, 1.0f / ms);
        ++M.step;
        const float bc1 = 1.0f - std::pow(b1, (float)M.step), bc2 = 1.0f - std::pow(b2, (float)M.step);
        hipLaunchKernelGGL(race::k_adam, dim3(race::blocks(race::kParams)), dim3(256), 0, M.st, M.p.p, M.g.p, M.m.p,
                           M.v.p, race::kParams, lr, b1, b2, eps, bc1, bc2);
    }
    RACE_HIP(hipGetLastError());
    float l = 0;
    RACE_HIP(hipMemcpyAsync(&l, M.loss.p, sizeof(float), hipMemcpyDeviceToHost, M.st));
    RACE_HIP(hipStreamSynchronize(M.st));
    if (loss) *loss = l;
    return race::OK;
}

int race_mlp_destroy(race_mlp h) {
    if (!h) return race::OK;
    hipStream_t st = h->m.st;
    delete h;
    if (st) (void)hipStreamDestroy(st);
    return race::OK;
}

}  // extern "C"
