// devmath.h -- device-side elementary functions and small DFTs for libmhit.
//
// Every routine here is evaluated as written (the library is built with
// -ffp-contract=off), so a host implementation of the same algorithm, with the
// same operation order, produces the same bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mhit {

// ln2 split Cody-Waite style: LN2_HI has its 21 low mantissa bits cleared, so
// e*LN2_HI is exact for |e| < 2^11; LN2_LO = ln2 - LN2_HI.
constexpr double LN2_HI = 6.93147180369123816490e-01;
constexpr double LN2_LO = 1.90821492927058770002e-10;

// Table of the table-driven log and expm1 (built on the host by
// build_math_table, identical to the oracle's):
//   [TAB_INVC + i]  1/c_i rounded to 8 significant bits, c_i = 1 + (i+1/2)/128,
//                   i < 128; entry 128 is 1 (the |x-1| < 2^-7 path);
//   [TAB_LHI + i], [TAB_LLO + i]   -log(invc_i) as a double-double;
//   [TAB_EHI + j], [TAB_ELO + j]   2^(j/32) as a double-double, j < 32.
constexpr int TAB_INVC = 0, TAB_LHI = 129, TAB_LLO = 258, TAB_EHI = 387, TAB_ELO = 419, MATH_TAB = 451;

// Natural logarithm.  x = 2^e m, m in [1,2); i = the top 7 fraction bits of
// m; r = m*invc_i - 1 is exact (invc_i has 8 significant bits, |r| <= 2^-7)
// and one fma forms it; log x = e ln2 - log(invc_i) + log1p(r), log1p by its
// Taylor series through r^8.  Within 2^-7 of 1, r = x - 1 (exact) and the
// table term is 0.  The leading sum is formed error-free (Knuth's TwoSum).
__device__ __forceinline__ double dlog_core(double x, int e, const double *__restrict__ tab)
{
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    int i = (int)((b >> 45) & 127);
    if (__builtin_fabs(x - 1.0) < 0.0078125) { i = 128; m = x; e = 0; }
    const double r = __builtin_fma(m, tab[TAB_INVC + i], -1.0);
    double P = -0.125;
    P = __builtin_fma(r, P, 1.0 / 7.0);
    P = __builtin_fma(r, P, -1.0 / 6.0);
    P = __builtin_fma(r, P, 0.2);
    P = __builtin_fma(r, P, -0.25);
    P = __builtin_fma(r, P, 1.0 / 3.0);
    P = __builtin_fma(r, P, -0.5);
    const double de = (double)e;
    const double hi = __builtin_fma(de, LN2_HI, tab[TAB_LHI + i]);
    const double lo = __builtin_fma(de, LN2_LO, tab[TAB_LLO + i]);
    const double t = hi + r;
    const double bb = t - hi;
    const double err = (hi - (t - bb)) + (r - bb);
    return t + (err + __builtin_fma(r * r, P, lo));
}

__device__ __forceinline__ double dlog(double x, const double *__restrict__ tab)
{
    int e = 0;
    if (__builtin_expect(!(__builtin_isnormal(x) && x > 0.0), 0)) {   // zero, negative, NaN, inf, subnormal
        if (!(x > 0.0)) return (x == 0.0) ? -__builtin_inf() : __builtin_nan("");
        if (__builtin_isinf(x)) return x;
        x = x * 18014398509481984.0;                                  // subnormal: scale by 2^54
        e = -54;
    }
    return dlog_core(x, e, tab);
}

// expm1(y) for y <= 0: y = k ln2/32 + r with |r| <= ln2/64 (Cody-Waite in two
// fmas), k = 32q + j, exp(y) = 2^q 2^(j/32) exp(r), expm1(r) by its Taylor
// series through r^8/8!; expm1(y) = (s - 1) + (s expm1(r) + s_lo) with
// s + s_lo = 2^q 2^(j/32).
__device__ __forceinline__ double dexpm1_neg(double y, const double *__restrict__ tab)
{
    if (y < -45.0) return -1.0;
    const double kd = __builtin_rint(y * 46.16624130844683);    // ~32/ln2 (only picks k)
    double r = __builtin_fma(-kd, LN2_HI / 32.0, y);
    r = __builtin_fma(-kd, LN2_LO / 32.0, r);
    double P = 1.0 / 40320.0;
    P = __builtin_fma(r, P, 1.0 / 5040.0);
    P = __builtin_fma(r, P, 1.0 / 720.0);
    P = __builtin_fma(r, P, 1.0 / 120.0);
    P = __builtin_fma(r, P, 1.0 / 24.0);
    P = __builtin_fma(r, P, 1.0 / 6.0);
    P = __builtin_fma(r, P, 0.5);
    const double p = __builtin_fma(r * r, P, r);
    const int k = (int)kd;
    if (k == 0) return p;
    const int j = k & 31, q = k >> 5;
    const double s = __builtin_ldexp(tab[TAB_EHI + j], q);
    const double slo = __builtin_ldexp(tab[TAB_ELO + j], q);
    return (s - 1.0) + __builtin_fma(s, p, slo);
}

// tanh(x) = -expm1(-2|x|) / (2 + expm1(-2|x|)), odd extension.
__device__ __forceinline__ double dtanh(double x, const double *__restrict__ tab)
{
    if (__builtin_isnan(x)) return x;
    double a = __builtin_fabs(x);
    double t;
    if (a < 3.725290298461914e-09) t = a;
    else if (a > 22.0) t = 1.0;
    else {
        double e = dexpm1_neg(-2.0 * a, tab);
        t = -e / (2.0 + e);
    }
    return (x < 0.0) ? -t : t;
}

// ---------------------------------------------------------------------------
// Small DFTs in natural order; SG = -1 forward, +1 inverse.
// ---------------------------------------------------------------------------
template <int SG>
__device__ __forceinline__ void dft2(double2 *v)
{
    double ar = v[0].x, ai = v[0].y, br = v[1].x, bi = v[1].y;
    v[0] = make_double2(ar + br, ai + bi);
    v[1] = make_double2(ar - br, ai - bi);
}

template <int SG>
__device__ __forceinline__ void dft3(double2 *v)
{
    const double S3 = 0.8660254037844386;
    double x0r = v[0].x, x0i = v[0].y, x1r = v[1].x, x1i = v[1].y, x2r = v[2].x, x2i = v[2].y;
    double tr = x1r + x2r, ti = x1i + x2i;
    double dr = x1r - x2r, di = x1i - x2i;
    double mr = x0r - 0.5 * tr, mi = x0i - 0.5 * ti;
    double pr = S3 * dr, pim = S3 * di;
    v[0] = make_double2(x0r + tr, x0i + ti);
    if (SG < 0) {
        v[1] = make_double2(mr + pim, mi - pr);
        v[2] = make_double2(mr - pim, mi + pr);
    } else {
        v[1] = make_double2(mr - pim, mi + pr);
        v[2] = make_double2(mr + pim, mi - pr);
    }
}

template <int SG>
__device__ __forceinline__ void dft4(double2 *v)
{
    double x0r = v[0].x, x0i = v[0].y, x1r = v[1].x, x1i = v[1].y;
    double x2r = v[2].x, x2i = v[2].y, x3r = v[3].x, x3i = v[3].y;
    double ar = x0r + x2r, ai = x0i + x2i, br = x0r - x2r, bi = x0i - x2i;
    double cr = x1r + x3r, ci = x1i + x3i, dr = x1r - x3r, di = x1i - x3i;
    v[0] = make_double2(ar + cr, ai + ci);
    v[2] = make_double2(ar - cr, ai - ci);
    if (SG < 0) {
        v[1] = make_double2(br + di, bi - dr);
        v[3] = make_double2(br - di, bi + dr);
    } else {
        v[1] = make_double2(br - di, bi + dr);
        v[3] = make_double2(br + di, bi - dr);
    }
}

template <int SG>
__device__ __forceinline__ void dft8(double2 *v)
{
    const double H = 0.7071067811865476;
    double2 e[4] = {v[0], v[2], v[4], v[6]};
    double2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4<SG>(e);
    dft4<SG>(o);
    double a, b;
    a = o[1].x; b = o[1].y;
    if (SG < 0) o[1] = make_double2((a + b) * H, (b - a) * H);
    else        o[1] = make_double2((a - b) * H, (a + b) * H);
    a = o[2].x; b = o[2].y;
    if (SG < 0) o[2] = make_double2(b, -a);
    else        o[2] = make_double2(-b, a);
    a = o[3].x; b = o[3].y;
    if (SG < 0) o[3] = make_double2((b - a) * H, -((a + b) * H));
    else        o[3] = make_double2(-((a + b) * H), (a - b) * H);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        v[q] = make_double2(e[q].x + o[q].x, e[q].y + o[q].y);
        v[q + 4] = make_double2(e[q].x - o[q].x, e[q].y - o[q].y);
    }
}

template <int R, int SG>
__device__ __forceinline__ void dft(double2 *v)
{
    if constexpr (R == 2) dft2<SG>(v);
    else if constexpr (R == 3) dft3<SG>(v);
    else if constexpr (R == 4) dft4<SG>(v);
    else dft8<SG>(v);
}

}  // namespace mhit
