// astro_internal.h -- helpers shared by the library's host sources: status
// reporting, device buffers, launch sizes.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/astro_abi.h"

namespace astro {

// Records msg as the calling thread's last error (astro_last_error) and
// returns code.
int fail(int code, const std::string& msg);

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::astro::fail(ASTRO_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define ASTRO_TRY(expr)                \
    do {                               \
        int s_ = (expr);               \
        if (s_ != ASTRO_OK) return s_; \
    } while (0)

// Device buffer owned by RAII.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return ASTRO_OK;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(ASTRO_ERR_DEVICE, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes: " +
                                              hipGetErrorString(e));
        }
        return ASTRO_OK;
    }
    int upload(const T* host, size_t count, hipStream_t s) {
        ASTRO_TRY(alloc(count));
        if (count) HIP_TRY(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
        return ASTRO_OK;
    }
    int upload(const std::vector<T>& v, hipStream_t s) { return upload(v.data(), v.size(), s); }
};

static inline int grid_for(int64_t n, int block, int cap = 65536) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace astro
