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

// Natural logarithm.  x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m - 1,
// s = f/(2+f); log(1+f) = 2 atanh(s) = 2s + s R(s^2) with the Taylor series
// R(z) = sum_{k=1..12} 2/(2k+1) z^k, evaluated as f - (f^2/2 - s(f^2/2 + R)).
__device__ __forceinline__ double dlog(double x)
{
    if (!(x > 0.0)) return (x == 0.0) ? -__builtin_inf() : __builtin_nan("");
    if (__builtin_isinf(x)) return x;
    int e = 0;
    if (x < 2.2250738585072014e-308) { x = x * 18014398509481984.0; e = -54; }
    uint64_t b = (uint64_t)__double_as_longlong(x);
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 2.0 / 25.0;
    R = 2.0 / 23.0 + z * R;
    R = 2.0 / 21.0 + z * R;
    R = 2.0 / 19.0 + z * R;
    R = 2.0 / 17.0 + z * R;
    R = 2.0 / 15.0 + z * R;
    R = 2.0 / 13.0 + z * R;
    R = 2.0 / 11.0 + z * R;
    R = 2.0 / 9.0 + z * R;
    R = 2.0 / 7.0 + z * R;
    R = 2.0 / 5.0 + z * R;
    R = 2.0 / 3.0 + z * R;
    R = z * R;
    double hfsq = 0.5 * (f * f);
    double r = f - (hfsq - s * (hfsq + R));
    double de = (double)e;
    return de * LN2_HI + (de * LN2_LO + r);
}

// expm1(y) for y <= 0: y = k ln2 + r with |r| <= ln2/2, expm1(r) by its Taylor
// series through r^15/15!, expm1(y) = 2^k p + (2^k - 1).
__device__ __forceinline__ double dexpm1_neg(double y)
{
    if (y < -45.0) return -1.0;
    double kd = __builtin_rint(y * 1.4426950408889634);
    double r = (y - kd * LN2_HI) - kd * LN2_LO;
    double p = 1.0 / 1307674368000.0;
    p = 1.0 / 87178291200.0 + r * p;
    p = 1.0 / 6227020800.0 + r * p;
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = r + (r * r) * p;
    int k = (int)kd;
    if (k == 0) return p;
    double two_k = __longlong_as_double((long long)((uint64_t)(k + 1023) << 52));
    return two_k * p + (two_k - 1.0);
}

// tanh(x) = -expm1(-2|x|) / (2 + expm1(-2|x|)), odd extension.
__device__ __forceinline__ double dtanh(double x)
{
    if (__builtin_isnan(x)) return x;
    double a = __builtin_fabs(x);
    double t;
    if (a < 3.725290298461914e-09) t = a;
    else if (a > 22.0) t = 1.0;
    else {
        double e = dexpm1_neg(-2.0 * a);
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
