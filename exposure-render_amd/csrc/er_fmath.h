/*
 * er_fmath.h - the float elementary functions of the tracer, written for this project.
 *
 * The reference calls libm (`log`, `expf`, `powf`, `sinf`, `cosf`, `acosf`, `atan2f`;
 * Source/raymarching.h:45,89, Source/tonemap.cuh:34-36, Source/shader.h:145,158,174,185,
 * Source/utilities.h:30, Source/sphere.h:157-158, Source/geometry.h:72,77). A device
 * libm and a host libm round differently, and a one-ulp difference in a sampled light
 * position can flip a free-path decision. These functions use only IEEE-754 single
 * precision +, -, *, /, floorf and integer bit manipulation, so with floating-point
 * contraction disabled (-ffp-contract=off) they return the same bits on gfx950 and on
 * x86-64. Methods:
 *   - sin/cos: Cody-Waite reduction by pi/2 (three-part constant; the first two parts
 *     have 12 significant bits so k*P1 and k*P2 are exact for |k| < 2^12), then Taylor
 *     polynomials of degree 9 (sin) and 10 (cos) on [-pi/4, pi/4].
 *   - exp: Cody-Waite reduction by ln 2 (two-part constant), Taylor polynomial of degree 8
 *     on |r| <= ln2/2, exact scaling by 2^k through the exponent field.
 *   - log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log(m) = 2 atanh(s), s = (m-1)/(m+1),
 *     odd series to s^13.
 *   - pow: exp(y log x) for x >= 0.
 *   - atan: reduction to |t| <= tan(pi/8) with atan(t) = pi/4 + atan((t-1)/(t+1)), odd
 *     Taylor series to t^21; atan2 and acos from it.
 * The polynomial coefficients are 1/n! and 1/n rounded to float. The tests measure the
 * error against double-precision libm (sin/cos/exp/log/pow within a few ulp on the domains
 * the tracer uses). Valid ranges: sin/cos |x| < 6000.
 */
#ifndef ER_FMATH_H
#define ER_FMATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define ER_HD __host__ __device__ inline
#else
#define ER_HD inline
#endif

namespace erm {

ER_HD float bits_to_float(uint32_t u) { return __builtin_bit_cast(float, u); }
ER_HD uint32_t float_to_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

/* 2^k for k in [-149, 127] built from the exponent field (exact). */
ER_HD float exp2i(int k)
{
    if (k >= -126) return bits_to_float((uint32_t)(k + 127) << 23);
    return bits_to_float(1u << (k + 149));   /* subnormal powers of two */
}

ER_HD float expf_(float x)
{
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97208404541015625f) return 0.0f;
    const float kf = floorf(x * 1.44269502162933349609375f + 0.5f);
    const int   k  = (int)kf;
    const float r  = (x - kf * 0.693115234375f) - kf * 3.19461832987144589424e-05f;
    float p = 2.4801587642286904e-05f;          /* 1/8! */
    p = p * r + 1.98412701138295233e-04f;        /* 1/7! */
    p = p * r + 1.38888892251998186e-03f;        /* 1/6! */
    p = p * r + 8.33333376795053482e-03f;        /* 1/5! */
    p = p * r + 4.16666679084300995e-02f;        /* 1/4! */
    p = p * r + 1.66666671633720398e-01f;        /* 1/3! */
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    /* scale by 2^k in two exact steps so that neither factor overflows nor underflows early */
    const int k1 = k / 2;
    const int k2 = k - k1;
    return (p * exp2i(k1)) * exp2i(k2);
}

/* exp(x) - 1 without the cancellation of expf_(x) - 1 near 0: for |x| < 1/4 the Taylor
 * series x + x^2 (1/2! + x/3! + ... + x^7/9!) (the first omitted term is below 3e-13 |x|);
 * beyond, expf_(x) - 1, whose cancellation costs at most a few ulp there. */
ER_HD float expm1f_(float x)
{
    if (x != x) return x;
    if (x > -0.25f && x < 0.25f) {
        float p = 2.75573188446287530e-06f;      /* 1/9! */
        p = p * x + 2.4801587642286904e-05f;     /* 1/8! */
        p = p * x + 1.98412701138295233e-04f;    /* 1/7! */
        p = p * x + 1.38888892251998186e-03f;    /* 1/6! */
        p = p * x + 8.33333376795053482e-03f;    /* 1/5! */
        p = p * x + 4.16666679084300995e-02f;    /* 1/4! */
        p = p * x + 1.66666671633720398e-01f;    /* 1/3! */
        p = p * x + 0.5f;
        return x + (x * x) * p;
    }
    return expf_(x) - 1.0f;
}

ER_HD float logf_(float x)
{
    if (x != x) return x;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }   /* subnormal: scale by 2^23 */
    uint32_t u = float_to_bits(x);
    e += (int)((u >> 23) & 0xff) - 127;
    uint32_t mb = (u & 0x007fffffu) | 0x3f800000u;              /* m in [1, 2) */
    if (mb > 0x3fb504f3u) { mb -= 0x00800000u; e += 1; }         /* m in [sqrt(1/2), sqrt(2)) */
    const float m  = bits_to_float(mb);
    const float f  = m - 1.0f;
    const float s  = f / (2.0f + f);
    const float s2 = s * s;
    float p = 7.69230797886848449707e-02f;       /* 1/13 */
    p = p * s2 + 9.09090936183929443359e-02f;    /* 1/11 */
    p = p * s2 + 1.11111111938953399658e-01f;    /* 1/9  */
    p = p * s2 + 1.42857149243354797363e-01f;    /* 1/7  */
    p = p * s2 + 2.00000002980232238770e-01f;    /* 1/5  */
    p = p * s2 + 3.33333343267440795898e-01f;    /* 1/3  */
    const float ef = (float)e;
    const float lm = (2.0f * s) + (2.0f * s) * (s2 * p);
    return ef * 0.693115234375f + (lm + ef * 3.19461832987144589424e-05f);
}

ER_HD float powf_(float x, float y)
{
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x != x || y != y) return x + y;
    if (x == 0.0f) return y > 0.0f ? 0.0f : __builtin_inff();
    if (x < 0.0f) return __builtin_nanf("");
    return expf_(y * logf_(x));
}

ER_HD void sincosf_(float x, float* s, float* c)
{
    const float kf = floorf(x * 0.636619746685028076171875f + 0.5f);
    const int   q  = ((int)kf) & 3;
    const float r  = ((x - kf * 1.5703125f) - kf * 4.83751296997070312500e-04f) - kf * 7.549790126404332e-08f;
    const float r2 = r * r;
    float ps = 2.75573188446287530e-06f;         /* 1/9! */
    ps = ps * r2 - 1.98412701138295233e-04f;     /* 1/7! */
    ps = ps * r2 + 8.33333376795053482e-03f;     /* 1/5! */
    ps = ps * r2 - 1.66666671633720398e-01f;     /* 1/3! */
    const float sr = r + (r * r2) * ps;
    float pc = -2.75573199814971303e-07f;        /* 1/10! */
    pc = pc * r2 + 2.48015876422869042e-05f;     /* 1/8! */
    pc = pc * r2 - 1.38888892251998186e-03f;     /* 1/6! */
    pc = pc * r2 + 4.16666679084300995e-02f;     /* 1/4! */
    pc = pc * r2 - 0.5f;
    const float cr = 1.0f + r2 * pc;
    switch (q) {
        case 0:  *s =  sr; *c =  cr; break;
        case 1:  *s =  cr; *c = -sr; break;
        case 2:  *s = -sr; *c = -cr; break;
        default: *s = -cr; *c =  sr; break;
    }
}

ER_HD float sinf_(float x) { float s, c; sincosf_(x, &s, &c); return s; }
ER_HD float cosf_(float x) { float s, c; sincosf_(x, &s, &c); return c; }

ER_HD float atanf_(float x)
{
    if (x != x) return x;
    const bool neg = x < 0.0f;
    float t = neg ? -x : x;
    const bool inv = t > 1.0f;
    if (inv) t = 1.0f / t;
    float base = 0.0f;
    if (t > 0.41421356237309503f) { t = (t - 1.0f) / (t + 1.0f); base = 0.785398185253143310546875f; }
    const float t2 = t * t;
    float p = 4.76190485060214996338e-02f;       /* 1/21 */
    p = p * t2 - 5.26315793395042419434e-02f;    /* 1/19 */
    p = p * t2 + 5.88235296308994293213e-02f;    /* 1/17 */
    p = p * t2 - 6.66666701436042785645e-02f;    /* 1/15 */
    p = p * t2 + 7.69230797886848449707e-02f;    /* 1/13 */
    p = p * t2 - 9.09090936183929443359e-02f;    /* 1/11 */
    p = p * t2 + 1.11111111938953399658e-01f;    /* 1/9  */
    p = p * t2 - 1.42857149243354797363e-01f;    /* 1/7  */
    p = p * t2 + 2.00000002980232238770e-01f;    /* 1/5  */
    p = p * t2 - 3.33333343267440795898e-01f;    /* 1/3  */
    float r = base + (t + (t * t2) * p);
    if (inv) r = 1.57079637050628662109375f - r;
    return neg ? -r : r;
}

ER_HD float atan2f_(float y, float x)
{
    const float pi = 3.1415927410125732421875f;
    if (x != x || y != y) return x + y;
    if (x == 0.0f) {
        if (y == 0.0f) return (float_to_bits(x) >> 31) ? ((float_to_bits(y) >> 31) ? -pi : pi) : y;
        return y > 0.0f ? 1.57079637050628662109375f : -1.57079637050628662109375f;
    }
    const float a = atanf_(y / x);
    if (x > 0.0f) return a;
    return (y >= 0.0f) ? a + pi : a - pi;
}

ER_HD float acosf_(float x)
{
    if (x != x) return x;
    if (x > 1.0f || x < -1.0f) return __builtin_nanf("");
    const float s = sqrtf((1.0f - x) * (1.0f + x));
    return atan2f_(s, x);
}

} /* namespace erm */

#endif /* ER_FMATH_H */
