// astro_cx.h -- complex double arithmetic for the device's complex-step
// kernels (design derivatives, co-rotational shell forces).
#pragma once
#include <hip/hip_runtime.h>

namespace astro {

struct cx {
    double r, i;
};
__device__ __forceinline__ cx mkc(double r, double i) { return cx{r, i}; }
__device__ __forceinline__ cx operator+(cx a, cx b) { return cx{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cx operator-(cx a, cx b) { return cx{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cx operator*(cx a, cx b) { return cx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cx operator*(double s, cx a) { return cx{s * a.r, s * a.i}; }
__device__ __forceinline__ cx operator*(cx a, double s) { return cx{a.r * s, a.i * s}; }
__device__ __forceinline__ cx operator+(cx a, double s) { return cx{a.r + s, a.i}; }
__device__ __forceinline__ cx operator/(cx a, cx b) {
    const double d = b.r * b.r + b.i * b.i;
    return cx{(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}
__device__ __forceinline__ cx operator/(double s, cx b) { return cx{s, 0.0} / b; }
__device__ __forceinline__ cx operator-(cx a) { return cx{-a.r, -a.i}; }

}  // namespace astro
