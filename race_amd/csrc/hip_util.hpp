// hip_util.hpp -- error checking and small device buffer helper.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstddef>

namespace race {

#define RACE_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) {                                                            \
            std::fprintf(stderr, "race_amd: HIP error %s at %s:%d (%s)\n",                 \
                         hipGetErrorString(_e), __FILE__, __LINE__, #call);                \
            return race::ERR_INTERNAL;                                                     \
        }                                                                                  \
    } while (0)

// Owning device allocation.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() {}
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // (re)allocate to at least `count` elements; contents are not kept
    hipError_t reserve(size_t count) {
        if (count <= n && p) return hipSuccess;
        release();
        size_t c = count ? count : 1;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), c * sizeof(T));
        if (e == hipSuccess) n = c;
        return e;
    }
};

}  // namespace race
