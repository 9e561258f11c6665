// cbc_math.h -- lean FP64 elementary functions for the CKM1D kernels.
//
// The generic device libm carries a Payne--Hanek path for arbitrarily large
// trig arguments and a double-double pow; inlined six times per Cash--Karp
// step they dominate the code size and register pressure of the stepper.  The
// trig arguments here are bounded: theta = 2*pi*(f_k/f_ref)*tau + phase, and
// the three-part reduction below is accurate for |theta| < 2^20*pi/2, i.e.
// horizons below ~2.5e5 periods of the highest harmonic (the Python layer
// checks time_end against this bound).  These versions use:
//
//   sincos  Cody & Waite (1980) argument reduction by pi/2 split in three
//           parts (FMA form), and the fdlibm __kernel_sin / __kernel_cos
//           minimax polynomials on [-pi/4, pi/4] (Sun Microsystems fdlibm
//           5.3, k_sin.c / k_cos.c coefficients S1..S6, C1..C6).
//   exp     reduction x = k*ln2 + r with a two-part ln2 (Cody--Waite), a
//           degree-13 Taylor polynomial of e^r on |r| <= ln2/2, ldexp.
//   log     x = 2^e * m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//           s = (m-1)/(m+1), odd series of atanh to s^23.
//   pow_pos x^y = exp(y * log x) for x > 0.
//   rcp     hardware reciprocal estimate plus two Newton--Raphson steps.
//
// Inside one Cash--Karp attempt the stage arguments stay close to those of
// the attempt's first stage, so the stages reuse the first stage's values:
//   sincos_kern  the fdlibm kernels alone (|r| <= pi/4), as sin r, cos r - 1;
//   sincos_tiny  Taylor series of sin b, cos b - 1 for |b| <= 1/20;
//   rotate       sin/cos(A + b) from sin/cos A by the addition theorem;
//   pow1p_small  (1 + e)^y = exp(y log1p e) by series for |e| <= 1/100, |y| <= 6.
//
// Accuracy against numpy float64 is checked on the device by
// cbc_selftest_math (tests/test_gpu_math.py): sincos, exp, log within 2 ulp,
// rcp within 1 ulp, pow_pos within (|y log x| + 4) ulp relative (the log's
// rounding scaled by the exponent).
#pragma once

#include <hip/hip_runtime.h>

namespace cbc {

// pi/2 = PIO2_1 + PIO2_2 + PIO2_3 (each the rounded remainder of the previous)
__device__ constexpr double kTwoOverPi = 6.36619772367581382433e-01;
__device__ constexpr double kPio2_1 = 1.57079632679489655800e+00;
__device__ constexpr double kPio2_2 = 6.12323399573676603587e-17;
__device__ constexpr double kPio2_3 = -1.4973849048591698e-33;

// 1/x from the hardware estimate v_rcp_f64 refined by two Newton--Raphson
// steps (e = 1 - x r, r += r e); within 1 ulp for normal |x| in
// [2^-1000, 2^1000], which covers every divisor of the model.
__device__ __forceinline__ double rcp_fast(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ void sincos_fast(double x, double* sp, double* cp)
{
    const double k = __builtin_rint(x * kTwoOverPi);
    double r = __builtin_fma(-k, kPio2_1, x);
    r = __builtin_fma(-k, kPio2_2, r);
    r = __builtin_fma(-k, kPio2_3, r);
    const int q = (int)k;

    // fdlibm k_sin.c
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    // fdlibm k_cos.c
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;

    const double z = r * r;
    const double v = z * r;
    const double ps = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    const double s = r + v * (S1 + z * ps);

    const double pc = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    const double c = w + (((1.0 - w) - hz) + z * pc);

    // quadrant: sin(r + q pi/2), cos(r + q pi/2)
    const bool swap = (q & 1) != 0;
    const double so = swap ? c : s;
    const double co = swap ? s : c;
    *sp = (q & 2) ? -so : so;
    *cp = ((q + 1) & 2) ? -co : co;
}

__device__ __forceinline__ double exp_fast(double x)
{
    const double kLog2e = 1.44269504088896338700e+00;
    const double kLn2Hi = 6.93147180369123816490e-01;  // high 32 bits of ln 2
    const double kLn2Lo = 1.90821492927058770002e-10;  // ln 2 - kLn2Hi
    const double k = __builtin_rint(x * kLog2e);
    const double r = __builtin_fma(-k, kLn2Lo, __builtin_fma(-k, kLn2Hi, x));
    // e^r = sum_{n<=13} r^n / n!
    double p = 1.0 / 6227020800.0;               // 1/13!
    p = __builtin_fma(p, r, 1.0 / 479001600.0);  // 1/12!
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)k);
}

__device__ __forceinline__ double log_fast(double x)
{
    int e;
    double m = __builtin_frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) * rcp_fast(m + 1.0);
    const double z = s * s;
    // atanh(s)/s = sum z^n / (2n+1), n = 0..11
    double p = 1.0 / 23.0;
    p = __builtin_fma(p, z, 1.0 / 21.0);
    p = __builtin_fma(p, z, 1.0 / 19.0);
    p = __builtin_fma(p, z, 1.0 / 17.0);
    p = __builtin_fma(p, z, 1.0 / 15.0);
    p = __builtin_fma(p, z, 1.0 / 13.0);
    p = __builtin_fma(p, z, 1.0 / 11.0);
    p = __builtin_fma(p, z, 1.0 / 9.0);
    p = __builtin_fma(p, z, 1.0 / 7.0);
    p = __builtin_fma(p, z, 1.0 / 5.0);
    p = __builtin_fma(p, z, 1.0 / 3.0);
    const double kLn2Hi = 6.93147180369123816490e-01;
    const double kLn2Lo = 1.90821492927058770002e-10;
    const double de = (double)e;
    // log m = 2 s + 2 s z p
    const double two_s = 2.0 * s;
    return __builtin_fma(de, kLn2Hi, __builtin_fma(two_s * z, p, __builtin_fma(de, kLn2Lo, two_s)));
}

// x^y for x > 0
__device__ __forceinline__ double pow_pos(double x, double y) { return exp_fast(y * log_fast(x)); }

constexpr double kPi4 = 0.78539816339744830962;  // bound of sincos_kern
constexpr double kTinyAngle = 0.05;              // bound of sincos_tiny
constexpr double kSmallRatio = 0.01;             // bound of pow1p_small (|e|)
constexpr double kSmallExp = 6.0;                // bound of pow1p_small (|y|)

// sin r and cos r - 1 for |r| <= pi/4: the fdlibm __kernel_sin / __kernel_cos
// polynomials of sincos_fast without the argument reduction.
__device__ __forceinline__ void sincos_kern(double r, double* s, double* cm1)
{
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    const double z = r * r;
    const double v = z * r;
    const double ps = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    *s = r + v * (S1 + z * ps);
    const double pc = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    *cm1 = z * pc - 0.5 * z;
}

// sin b and cos b - 1 for |b| <= 1/20 by Taylor series to b^7 and b^8; the
// truncation is below 2^-62 relative.
__device__ __forceinline__ void sincos_tiny(double b, double* s, double* cm1)
{
    const double z = b * b;
    *s = b + (b * z) * (-1.0 / 6.0 + z * (1.0 / 120.0 + z * (-1.0 / 5040.0)));
    *cm1 = z * (-0.5 + z * (1.0 / 24.0 + z * (-1.0 / 720.0 + z * (1.0 / 40320.0))));
}

// sin/cos(A + b) = sin A + (sin A (cos b - 1) + cos A sin b),
//                  cos A + (cos A (cos b - 1) - sin A sin b)
__device__ __forceinline__ void rotate(double sa, double ca, double sb, double cm1, double* so, double* co)
{
    *so = sa + __builtin_fma(ca, sb, sa * cm1);
    *co = ca + __builtin_fma(-sa, sb, ca * cm1);
}

// (1 + e)^y for |e| <= 1/100 and |y| <= 6: log1p by its series to e^8, exp of
// |y log1p e| <= 0.061 by Taylor to degree 10; truncation below 2^-60 relative.
__device__ __forceinline__ double pow1p_small(double e, double y)
{
    double l = -1.0 / 8.0;
    l = __builtin_fma(l, e, 1.0 / 7.0);
    l = __builtin_fma(l, e, -1.0 / 6.0);
    l = __builtin_fma(l, e, 1.0 / 5.0);
    l = __builtin_fma(l, e, -1.0 / 4.0);
    l = __builtin_fma(l, e, 1.0 / 3.0);
    l = __builtin_fma(l, e, -1.0 / 2.0);
    l = __builtin_fma(l, e, 1.0);
    const double u = y * (l * e);
    double p = 1.0 / 3628800.0;                 // 1/10!
    p = __builtin_fma(p, u, 1.0 / 362880.0);
    p = __builtin_fma(p, u, 1.0 / 40320.0);
    p = __builtin_fma(p, u, 1.0 / 5040.0);
    p = __builtin_fma(p, u, 1.0 / 720.0);
    p = __builtin_fma(p, u, 1.0 / 120.0);
    p = __builtin_fma(p, u, 1.0 / 24.0);
    p = __builtin_fma(p, u, 1.0 / 6.0);
    p = __builtin_fma(p, u, 0.5);
    p = __builtin_fma(p, u, 1.0);
    return __builtin_fma(p, u, 1.0);
}

}  // namespace cbc
