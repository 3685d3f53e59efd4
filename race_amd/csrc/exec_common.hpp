// exec_common.hpp -- helpers shared by the kernel files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "engine.hpp"
#include "hip_util.hpp"

namespace race {

constexpr int kBlock = 256;
constexpr size_t kMaxLds = 64 * 1024;  // LDS bytes one workgroup may allocate

template <int W>
__device__ __forceinline__ double groupSum(double v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// bfloat16 bits of v (through float, round to nearest even)
inline uint16_t toBf16(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
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

// lanes per row from the mean row length (row-parallel CRS kernels)
inline int pickTpr(long nnz, int n) {
    double m = n ? (double)nnz / n : 1;
    if (m <= 6) return 4;
    if (m <= 12) return 8;
    if (m <= 24) return 16;
    return 32;
}

}  // namespace race
