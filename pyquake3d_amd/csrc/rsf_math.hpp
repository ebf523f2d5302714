// Elementary functions for the rate-and-state right-hand side, on the domain
// its callers guarantee (host and device; the host build serves the accuracy
// test, tests/test_rsf_math_cpu.py).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RSF_HD __host__ __device__ __forceinline__
#else
#define RSF_HD inline
#endif

#include <cmath>

namespace pq3d {

// exp(x) for |x| <= 708.  Every caller clips its argument to +-700 first, so
// no overflow, underflow or NaN handling is needed.  Cody-Waite reduction
// x = k ln2 + r, |r| <= ln2 / 2, with ln2 split into a head of 32 significant
// bits and a tail; exp(r) by its Taylor series to degree 13 (the truncation
// is below 5e-18 relative at |r| = ln2 / 2), in Horner form; 2^k by ldexp.
RSF_HD double exp_clipped(double x) {
    const double k = std::rint(x * 1.4426950408889634);
    double r = std::fma(-k, 6.93147180369123816490e-01, x);
    r = std::fma(-k, 1.90821492927058770002e-10, r);
    double p = 1.0 / 6227020800.0;
    p = std::fma(p, r, 1.0 / 479001600.0);
    p = std::fma(p, r, 1.0 / 39916800.0);
    p = std::fma(p, r, 1.0 / 3628800.0);
    p = std::fma(p, r, 1.0 / 362880.0);
    p = std::fma(p, r, 1.0 / 40320.0);
    p = std::fma(p, r, 1.0 / 5040.0);
    p = std::fma(p, r, 1.0 / 720.0);
    p = std::fma(p, r, 1.0 / 120.0);
    p = std::fma(p, r, 1.0 / 24.0);
    p = std::fma(p, r, 1.0 / 6.0);
    p = std::fma(p, r, 0.5);
    p = std::fma(p, r, 1.0);
    p = std::fma(p, r, 1.0);
    return std::ldexp(p, (int)k);
}

// 1 / x for finite, nonzero, normal x: on the device the hardware estimate
// refined by two Newton steps (within an ulp of the quotient, without the
// IEEE division's scaling and fix-up); on the host the division itself.
RSF_HD double rcp_fast(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = std::fma(-x, r, 1.0);
    r = std::fma(r, e, r);
    e = std::fma(-x, r, 1.0);
    return std::fma(r, e, r);
#else
    return 1.0 / x;
#endif
}

}  // namespace pq3d
