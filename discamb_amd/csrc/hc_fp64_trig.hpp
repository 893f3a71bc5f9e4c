// FP64 phase factor and Debye-Waller factor of the FP64 kernels, written for their argument
// ranges: the phase arrives already reduced to [0, 1) revolutions and the Debye-Waller exponent
// is non-negative, so neither needs a general-argument reduction or special-value handling.
//
// sincos_rev: r = t - rint(t) in [-1/2, 1/2], quadrant q = rint(4 r), s = r - q/4 in [-1/8, 1/8]
// (exact), then the Taylor series of sin(2 pi s) to s^15 and cos(2 pi s) to s^16 (truncation
// below 5e-17) and the quadrant rotation.  exp_neg: e log2(e) = n + f with |f| <= 1/2, 2^-f by the
// Taylor series of exp(-f ln 2) to degree 12 (truncation below 2e-16), scaled by 2^-n.
// Coefficients: (-1)^k (2 pi)^(2k+1) / (2k+1)!, (-1)^k (2 pi)^(2k) / (2k)!, 1 / k!.
#pragma once

#include <hip/hip_runtime.h>

namespace discamb_amd {

__host__ __device__ __forceinline__ void sincos_rev(double t, double* sn, double* cs)
{
    const double r = t - rint(t);
    const double q = rint(4.0 * r);
    const double s = fma(-0.25, q, r);
    const double z = s * s;
    double p = -7.1812230177850051223e-1;
    p = fma(p, z, 3.8199525848482821277);
    p = fma(p, z, -1.5094642576822990392e+1);
    p = fma(p, z, 4.2058693944897653145e+1);
    p = fma(p, z, -7.6705859753061385842e+1);
    p = fma(p, z, 8.1605249276075054203e+1);
    p = fma(p, z, -4.1341702240399760234e+1);
    p = fma(p, z, 6.2831853071795864769);
    const double sv = s * p;
    double c = 2.8200596845579121507e-1;
    c = fma(c, z, -1.7143907110886720654);
    c = fma(c, z, 7.9035363713184688042);
    c = fma(c, z, -2.6426256783374397453e+1);
    c = fma(c, z, 6.0244641371876660363e+1);
    c = fma(c, z, -8.5456817206693727736e+1);
    c = fma(c, z, 6.4939394022668291491e+1);
    c = fma(c, z, -1.9739208802178717238e+1);
    c = fma(c, z, 1.0);
    // (sin, cos)(2 pi (s + q/4)) for q mod 4 = 0: (S, C), 1: (C, -S), 2: (-S, -C), 3: (-C, S)
    const int qi = (int)q & 3;
    const double a = (qi & 1) ? c : sv, b = (qi & 1) ? sv : c;
    *sn = (qi & 2) ? -a : a;
    *cs = ((qi + 1) & 2) ? -b : b;
}

__host__ __device__ __forceinline__ double exp_neg(double e)
{
    const double y = fmin(fmax(e * 1.4426950408889634074, -1000.0), 2000.0);   // 2^-2000 flushes to 0
    const double n = rint(y);
    const double x = (n - y) * 6.9314718055994530942e-1;
    double p = 1.0 / 479001600.0;
    p = fma(p, x, 1.0 / 39916800.0);
    p = fma(p, x, 1.0 / 3628800.0);
    p = fma(p, x, 1.0 / 362880.0);
    p = fma(p, x, 1.0 / 40320.0);
    p = fma(p, x, 1.0 / 5040.0);
    p = fma(p, x, 1.0 / 720.0);
    p = fma(p, x, 1.0 / 120.0);
    p = fma(p, x, 1.0 / 24.0);
    p = fma(p, x, 1.0 / 6.0);
    p = fma(p, x, 0.5);
    p = fma(p, x, 1.0);
    p = fma(p, x, 1.0);
    return ldexp(p, -(int)n);
}

}  // namespace discamb_amd
