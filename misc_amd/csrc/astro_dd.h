// astro_dd.h -- double-double helpers shared by the kernel sources.
#pragma once
#include <hip/hip_runtime.h>

namespace astro {

// Double-double arithmetic (value = hi + lo, |lo| <= ulp(hi) / 2) for true
// residuals: TwoSum (Knuth) and TwoProduct through fma give the rounding
// error of each sum and product exactly.  Both need every product and sum
// rounded on its own: contraction (which would fuse a*b into the sum that
// follows, s = fma(a, b, hi)) is switched off in these helpers.
__device__ __forceinline__ void dd_two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void dd_add(double& hi, double& lo, double bh, double bl) {
#pragma clang fp contract(off)
    double s, e;
    dd_two_sum(hi, bh, s, e);
    e += lo + bl;
    hi = s + e;
    lo = e - (hi - s);
}
__device__ __forceinline__ void dd_add_prod(double& hi, double& lo, double a, double b) {
#pragma clang fp contract(off)
    const double p = a * b;
    dd_add(hi, lo, p, __builtin_fma(a, b, -p));
}

}  // namespace astro
