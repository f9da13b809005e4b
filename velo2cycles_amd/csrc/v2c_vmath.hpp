// v2c_vmath.hpp -- FP64 quotient, logarithm and exponential for the
// viscosity iterations.
//
// The viscosity solves (the quasi-Newton of ductilevisc and the per-marker
// Newton of the local update) spend most of their instructions in divisions
// and logarithms of well-scaled positive values (viscosities 1e16..1e30,
// stresses, their sums and ratios).  The general-purpose forms handle
// denormal, overflowing and special operands with extra scaling and
// double-double steps that these values never need:
//
//  * quot(a, b): the hardware reciprocal estimate refined by two
//    Newton-Raphson steps, then one residual correction of the quotient
//    (Markstein's method); within an ulp of a/b, almost always the correctly
//    rounded quotient.  b must be a normal double with 2^-1020 < |b| < 2^1020.
//  * log_pos(x): ln x = e ln 2 + 2 atanh(s), x = 2^e m, m in [sqrt(1/2),
//    sqrt(2)), s = (m - 1)/(m + 1), |s| <= 0.1716; the Taylor series of atanh
//    through s^21 (truncation below 2^-60 relative) and ln 2 split into a
//    21-bit head, exact in e ln 2 for any exponent, and its tail (Cody and
//    Waite).  About 2 ulp near x = 1, 1 ulp elsewhere; NaN outside
//    [2^-1020, 2^1020].
//  * exp_red(x): e^x = 2^k e^r, k = rint(x / ln 2), r = x - k ln 2 with ln 2
//    split into a 42-bit head (k ln 2_head exact) and its tail (Cody and
//    Waite), |r| <= 0.347, and the Taylor series of e^r through r^13/13!
//    (truncation below 2^-57 relative).  About an ulp; NaN for |x| >= 708.
//
// Arguments outside these ranges give NaN rather than a library fallback, so
// that no general-purpose code (and its constants) sits inside the callers'
// loops: a caller whose result turns NaN recomputes it with the library
// functions.
//
// The polynomial coefficients are pinned to scalar registers (sconst), so
// every Horner step is one FMA taking its coefficient as the scalar operand.
//
// All are host-callable (the host uses the IEEE reciprocal), so the same code
// is checked on the CPU against long-double references.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace v2c {

// A constant held in a scalar register pair (the empty asm only fixes the
// register class; it emits no instruction).
__host__ __device__ __forceinline__ double sconst(double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+s"(c));
#endif
    return c;
}

__host__ __device__ __forceinline__ double rcp_refined(double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / b;
#endif
}

__host__ __device__ __forceinline__ double quot(double a, double b)
{
    const double r = rcp_refined(b);
    const double y = a * r;
    const double e = fma(-b, y, a);
    return fma(e, r, y);
}

__host__ __device__ __forceinline__ double log_pos(double x)
{
    const bool ok = x >= 0x1p-1020 && x <= 0x1p1020;
    int e;
#if defined(__HIP_DEVICE_COMPILE__)
    double m = __builtin_amdgcn_frexp_mant(x);
    e = __builtin_amdgcn_frexp_exp(x);
#else
    double m = frexp(x, &e);
#endif
    if (m < 0.70710678118654752440) {   // m in [0.5, 1): move to [sqrt(1/2), sqrt(2))
        m = m * 2.0;
        e -= 1;
    }
    const double f = m - 1.0;           // exact
    const double s = quot(f, 2.0 + f);
    const double z = s * s;
    double p = sconst(1.0 / 21);
    p = fma(p, z, sconst(1.0 / 19));
    p = fma(p, z, sconst(1.0 / 17));
    p = fma(p, z, sconst(1.0 / 15));
    p = fma(p, z, sconst(1.0 / 13));
    p = fma(p, z, sconst(1.0 / 11));
    p = fma(p, z, sconst(1.0 / 9));
    p = fma(p, z, sconst(1.0 / 7));
    p = fma(p, z, sconst(1.0 / 5));
    p = fma(p, z, sconst(1.0 / 3));
    const double s2 = s + s;
    const double lnm = fma(s2, p * z, s2);                 // 2 atanh(s)
    const double de = (double)e;
    const double ln2_hi = sconst(0x1.62e42p-1), ln2_lo = sconst(0x1.fdf473de6af28p-22);
    const double r = fma(de, ln2_hi, fma(de, ln2_lo, lnm));
    return ok ? r : __builtin_nan("");
}

__host__ __device__ __forceinline__ double exp_red(double x)
{
    const bool ok = fabs(x) < 708.0;
    const double k = ok ? rint(x * sconst(0x1.71547652b82fep+0)) : 0.0;
    double r = fma(-k, sconst(0x1.62e42fefa38p-1), x);
    r = fma(-k, sconst(0x1.ef35793c7673p-45), r);
    double p = sconst(1.0 / 6227020800.0);
    p = fma(p, r, sconst(1.0 / 479001600.0));
    p = fma(p, r, sconst(1.0 / 39916800.0));
    p = fma(p, r, sconst(1.0 / 3628800.0));
    p = fma(p, r, sconst(1.0 / 362880.0));
    p = fma(p, r, sconst(1.0 / 40320.0));
    p = fma(p, r, sconst(1.0 / 5040.0));
    p = fma(p, r, sconst(1.0 / 720.0));
    p = fma(p, r, sconst(1.0 / 120.0));
    p = fma(p, r, sconst(1.0 / 24.0));
    p = fma(p, r, sconst(1.0 / 6.0));
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ok ? ldexp(p, (int)k) : __builtin_nan("");
}

// x^y for x > 0 as exp(y ln x), for the power laws of the viscosity (relative
// error a few 1e-15 at the exponents that occur, |y ln x| < 200).
__host__ __device__ __forceinline__ double pow_pos(double x, double y)
{
    return exp_red(y * log_pos(x));
}

}  // namespace v2c
