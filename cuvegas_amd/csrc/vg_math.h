// Deterministic FP64 log / exp / pow for the map-refinement and reallocation
// kernels.
//
// The stratification weight d_h = (sigma_h^2)^beta (reference
// cuda_wrapper.cu:41-47) feeds the integer allocation
// n_h = max(2, (int)(d_h / D * N)) (cuda_wrapper.cu:126-130), which is
// discontinuous: one ulp of difference in d_h can move n_h by one.  Device and
// host maths libraries may differ by an ulp, so these kernels use this routine
// written from two textbook methods, with every multiply-add an explicit fma:
//   * log: x = m * 2^e with m in [sqrt(1/2), sqrt(2)); with s = (m-1)/(m+1),
//     log m = 2 * atanh(s) = 2 (s + s^3/3 + s^5/5 + ...), |s| <= 0.1716, summed
//     to the s^23 term (truncation < 1e-18 relative); e*ln2 is added with ln2
//     split into a 32-bit head and a tail (Cody & Waite, "Software Manual for
//     the Elementary Functions", 1980).
//   * exp: y = k ln2 + r with the same split, |r| <= ln2/2, exp(r) by its
//     Taylor series to r^13/13! (truncation < 5e-18, Estrin's evaluation
//     order), then scaled by 2^k.
//   * pow(x, a) = exp(a log x) for x >= 0.
// Relative error is a few ulp for log/exp and below 1e-13 for pow over the
// range used; the oracle restates the same steps, so results agree bit for bit.
#pragma once

#ifndef VG_HD
#define VG_HD __host__ __device__ inline
#endif

#define VG_LN2_HI 6.93147180369123816490e-01  /* 0x3fe62e42fee00000 */
#define VG_LN2_LO 1.90821492927058770002e-10  /* 0x3dea39ef35793c76 */
#define VG_INV_LN2 1.44269504088896338700e+00
#define VG_SQRT2 1.41421356237309514547e+00

VG_HD double vg_from_bits(unsigned long long b) { return __builtin_bit_cast(double, b); }
VG_HD unsigned long long vg_to_bits(double x) { return __builtin_bit_cast(unsigned long long, x); }

VG_HD double vg_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return vg_from_bits(0x7ff8000000000000ull);
    if (x == 0.0) return -vg_from_bits(0x7ff0000000000000ull);
    if (x == vg_from_bits(0x7ff0000000000000ull)) return x;
    int adj = 0;
    if (x < 2.2250738585072014e-308) { x *= 18014398509481984.0; adj = -54; }  // 2^54
    const unsigned long long b = vg_to_bits(x);
    int e = (int)((b >> 52) & 0x7ffull) - 1023 + adj;
    double m = vg_from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > VG_SQRT2) { m *= 0.5; e += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double logm = fma((s + s) * z, p, s + s);
    const double de = (double)e;
    return fma(de, VG_LN2_HI, fma(de, VG_LN2_LO, logm));
}

// The argument is clamped to [-746, 710], where 2^k still lands in ldexp's
// underflow-to-0 / overflow-to-inf range, and the result is scaled by one
// ldexp (exact in the normal range, one rounding below it): no branch but the
// NaN test.
VG_HD double vg_exp(double y) {
    if (y != y) return y;
    const double yc = y < -746.0 ? -746.0 : (y > 710.0 ? 710.0 : y);
    const double kd = floor(fma(yc, VG_INV_LN2, 0.5));
    const double r = fma(-kd, VG_LN2_LO, fma(-kd, VG_LN2_HI, yc));
    // Taylor polynomial sum_{i<=13} r^i / i!: the r^4..r^13 tail by Estrin's
    // scheme (pairs, then r^2, r^4 levels), the four leading terms by Horner
    // steps last, where the rounding of the large terms matters most.
    const double r2 = r * r, r4 = r2 * r2;
    const double q0 = fma(r, 1.0 / 120.0, 1.0 / 24.0);
    const double q1 = fma(r, 1.0 / 5040.0, 1.0 / 720.0);
    const double q2 = fma(r, 1.0 / 362880.0, 1.0 / 40320.0);
    const double q3 = fma(r, 1.0 / 39916800.0, 1.0 / 3628800.0);
    const double q4 = fma(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);
    const double s0 = fma(q1, r2, q0);
    const double s1 = fma(q3, r2, q2);
    const double t = fma(fma(q4, r4, s1), r4, s0);
    double p = fma(t, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)kd);
}

// x^a for x >= 0 (the only case the integrator needs).
VG_HD double vg_pow(double x, double a) {
    if (x != x || a != a || x < 0.0) return vg_from_bits(0x7ff8000000000000ull);
    if (a == 0.0 || x == 1.0) return 1.0;
    const double inf = vg_from_bits(0x7ff0000000000000ull);
    if (x == 0.0) return a > 0.0 ? 0.0 : inf;
    if (x == inf) return a > 0.0 ? inf : 0.0;
    return vg_exp(a * vg_log(x));
}

// C-style (int) conversion with the saturating semantics of a hardware
// truncating convert: NaN -> 0, out-of-range values clamp to the int range.
VG_HD int vg_sat_int(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (-2147483647 - 1);
    return (int)v;
}
