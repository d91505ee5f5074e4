// Scalar FP64 helpers shared by the kernels and the host tests.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
// Plain host compilers (the CPU tests build this header with g++).
#define __host__
#define __device__
#define __forceinline__ inline
#endif

#include <cmath>

namespace lpg {

// Natural logarithm of a finite x > 0 (NaN for x <= 0 or a non-finite x), within about 2 ulp.
//
// Method: the textbook argument reduction x = 2^e m with m in [1/sqrt 2, sqrt 2) and the series
//   log m = 2 atanh s = 2 s (1 + s^2/3 + s^4/5 + ...),  s = (m - 1) / (m + 1),  |s| <= 3 - 2 sqrt 2 < 0.1716,
// truncated after the s^18 / 19 term (the first omitted term is below 2^-56 of the sum), then
// log x = e ln 2 + log m with ln 2 split into a 32-bit head (e * head is exact for |e| < 2^21) and a tail.
// m - 1 is exact (Sterbenz); 1 / (m + 1) comes from the hardware reciprocal estimate on the device (refined by two
// Newton steps) and from a division on the host.  About 35 instructions against twice that for the library's
// correctly rounded routine; the callers need the logarithm of J to a few ulp (the energy, gradient and Hessian
// are compared at 1e-12).
__host__ __device__ __forceinline__ double log_pos(double x) {
    int e = 0;
    double m = frexp(x, &e); // m in [0.5, 1)
    const bool low = m < 0.70710678118654752440;
    m = low ? m + m : m;
    e = low ? e - 1 : e;
    const double f = m - 1.0;
    const double d = m + 1.0;
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(d);
    double t = fma(-d, r, 1.0);
    r = fma(r, t, r);
    t = fma(-d, r, 1.0);
    r = fma(r, t, r);
    const double s = f * r;
#else
    const double s = f / d;
#endif
    const double z = s * s;
    // sum_{k=1..9} z^(k-1) * 2 / (2k + 1), Horner from the highest term.
    double P = 2.0 / 19.0;
    P = fma(P, z, 2.0 / 17.0);
    P = fma(P, z, 2.0 / 15.0);
    P = fma(P, z, 2.0 / 13.0);
    P = fma(P, z, 2.0 / 11.0);
    P = fma(P, z, 2.0 / 9.0);
    P = fma(P, z, 2.0 / 7.0);
    P = fma(P, z, 2.0 / 5.0);
    P = fma(P, z, 2.0 / 3.0);
    const double logm = fma(s * z, P, s + s);
    constexpr double kLn2Head = 0.69314718060195446014404296875; // ln 2 rounded to 32 significant bits (exact)
    constexpr double kLn2Tail = -4.2009150726810846e-11;         // ln 2 - kLn2Head, rounded
    const double ed = static_cast<double>(e);
    const double y = fma(ed, kLn2Head, fma(ed, kLn2Tail, logm));
    return (x > 0.0 && x <= 1.79769313486231570815e308) ? y : NAN;
}

} // namespace lpg
