/*
 * omath.h -- portable single-precision transcendentals for the oracle.
 *
 * TEST INFRASTRUCTURE: part of the CPU oracle; nothing in the product links it.
 *
 * The reference called CUDA's fast-math intrinsics (noor/noor.vcxproj:320),
 * whose results cannot be reproduced without NVIDIA hardware.  The oracle and
 * the HIP kernels therefore both use the same published algorithms, written
 * out in full on each side with identical operation order, so that CPU and
 * GPU produce bit-identical results:
 *   - sin / cos / atan / asin / exp / log: Cody-Waite argument reduction and
 *     the single-precision minimax polynomials of S. L. Moshier's Cephes
 *     library ("Methods and Programs for Mathematical Functions", 1989).
 *   - acos via the half-angle identity acos(x) = 2 asin(sqrt((1-x)/2)).
 *   - atan2 from atan with the usual quadrant fix-up.
 *   - pow(x, y) = exp(y log x) for x > 0.
 * Polynomials are evaluated with fmaf (correctly rounded on both sides).
 * Domain: |x| <= 8192 for sin/cos (accuracy degrades beyond).
 * Accuracy is checked against libm double precision in
 * tests/test_oracle_math.py.
 */
#ifndef NOOR_ORACLE_OMATH_H
#define NOOR_ORACLE_OMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#define OM_PI      3.14159265358979323846f
#define OM_PI_2    1.57079632679489661923f
#define OM_PI_4    0.785398163397448309616f

static inline uint32_t om_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float om_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* x * 2^n for n in [-252, 254] with a single final rounding. */
static inline float om_ldexp(float x, int n) {
    int n1 = n / 2;
    int n2 = n - n1;
    /* both halves are in [-126, 127]: exact normal powers of two */
    float s1 = om_u2f((uint32_t)(n1 + 127) << 23);
    float s2 = om_u2f((uint32_t)(n2 + 127) << 23);
    return (x * s1) * s2;
}

/* sin and cos on [-pi/4, pi/4] (Cephes sinf.c / cosf.c coefficients) */
static inline float om_sin_poly(float x) {
    float z = x * x;
    float p = fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f);
    p = fmaf(p, z, -1.6666654611e-1f);
    return fmaf(p * z, x, x);
}
static inline float om_cos_poly(float x) {
    float z = x * x;
    float p = fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f);
    p = fmaf(p, z, 4.166664568298827e-2f);
    return fmaf(p * z, z, fmaf(-0.5f, z, 1.0f));
}

/* Reduce x to r in [-pi/4, pi/4] and return the octant index j (even). */
static inline int om_reduce(float x, float* r) {
    float ax = fabsf(x);
    int j = (int)(ax * 1.27323954473516f);   /* 4/pi */
    j += (j & 1);
    float y = (float)j;
    /* pi/4 split in three parts (Cephes DP1..DP3) */
    float t = fmaf(-y, 0.78515625f, ax);
    t = fmaf(-y, 2.4187564849853515625e-4f, t);
    t = fmaf(-y, 3.77489497744594108e-8f, t);
    *r = t;
    return j;
}

static inline float om_sinf(float x) {
    if (!(fabsf(x) <= 3.0e38f)) return x - x;  /* nan, inf -> nan */
    float r;
    int j = om_reduce(x, &r);
    int q = (j >> 1) & 3;
    float v = (q & 1) ? om_cos_poly(r) : om_sin_poly(r);
    if (q & 2) v = -v;
    return (x < 0.0f) ? -v : v;
}

static inline float om_cosf(float x) {
    if (!(fabsf(x) <= 3.0e38f)) return x - x;
    float r;
    int j = om_reduce(x, &r);
    int q = (j >> 1) & 3;
    float v = (q & 1) ? om_sin_poly(r) : om_cos_poly(r);
    if (q == 1 || q == 2) v = -v;
    return v;
}

/* atan (Cephes atanf.c) */
static inline float om_atanf(float x) {
    float sign = 1.0f;
    if (x < 0.0f) { sign = -1.0f; x = -x; }
    float y;
    if (x > 2.414213562373095f) {        /* tan(3pi/8) */
        y = OM_PI_2;
        x = -1.0f / x;
    } else if (x > 0.4142135623730950f) { /* tan(pi/8) */
        y = OM_PI_4;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    float z = x * x;
    float p = fmaf(8.05374449538e-2f, z, -1.38776856032e-1f);
    p = fmaf(p, z, 1.99777106478e-1f);
    p = fmaf(p, z, -3.33329491539e-1f);
    y = y + fmaf(p * z, x, x);
    return sign * y;
}

static inline float om_atan2f(float y, float x) {
    if (x != x || y != y) return x + y;
    if (x == 0.0f) {
        if (y > 0.0f) return OM_PI_2;
        if (y < 0.0f) return -OM_PI_2;
        return 0.0f;
    }
    if (y == 0.0f) return (x > 0.0f) ? 0.0f : OM_PI;
    float w = 0.0f;
    if (x < 0.0f) w = (y < 0.0f) ? -OM_PI : OM_PI;
    return w + om_atanf(y / x);
}

/* asin on [-1, 1] (Cephes asinf.c) */
static inline float om_asinf(float x) {
    float a = fabsf(x);
    if (a > 1.0f) return (x - x) / (x - x);
    float z, t;
    int flag = 0;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        t = sqrtf(z);
        flag = 1;
    } else {
        t = a;
        z = a * a;
    }
    float p = fmaf(4.2163199048e-2f, z, 2.4181311049e-2f);
    p = fmaf(p, z, 4.5470025998e-2f);
    p = fmaf(p, z, 7.4953002686e-2f);
    p = fmaf(p, z, 1.6666752422e-1f);
    float r = fmaf(p * z, t, t);
    if (flag) r = OM_PI_2 - (r + r);
    return (x < 0.0f) ? -r : r;
}

static inline float om_acosf(float x) {
    if (x != x || fabsf(x) > 1.0f) return (x - x) / (x - x);
    if (x < -0.5f) return OM_PI - 2.0f * om_asinf(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * om_asinf(sqrtf(0.5f * (1.0f - x)));
    return OM_PI_2 - om_asinf(x);
}

/* exp (Cephes expf.c) */
static inline float om_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283905206835f) return INFINITY;
    if (x < -103.278929903431851103f) return 0.0f;
    float z = floorf(fmaf(1.44269504088896341f, x, 0.5f));
    int n = (int)z;
    x = fmaf(-z, 0.693359375f, x);
    x = fmaf(-z, -2.12194440e-4f, x);
    float p = fmaf(1.9875691500e-4f, x, 1.3981999507e-3f);
    p = fmaf(p, x, 8.3334519073e-3f);
    p = fmaf(p, x, 4.1665795894e-2f);
    p = fmaf(p, x, 1.6666665459e-1f);
    p = fmaf(p, x, 5.0000001201e-1f);
    float xx = x * x;
    float r = fmaf(p, xx, x) + 1.0f;
    return om_ldexp(r, n);
}

/* log (Cephes logf.c) */
static inline float om_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return (x - x) / (x - x);
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e = 0;
    uint32_t u = om_f2u(x);
    if ((u >> 23) == 0) {            /* subnormal: scale by 2^25 */
        x = x * 33554432.0f;
        u = om_f2u(x);
        e = -25;
    }
    e += (int)(u >> 23) - 126;
    x = om_u2f((u & 0x007FFFFFu) | 0x3F000000u);   /* mantissa in [0.5, 1) */
    if (x < 0.707106781186547524f) {
        e -= 1;
        x = x + x - 1.0f;
    } else {
        x = x - 1.0f;
    }
    float z = x * x;
    float p = fmaf(7.0376836292e-2f, x, -1.1514610310e-1f);
    p = fmaf(p, x, 1.1676998740e-1f);
    p = fmaf(p, x, -1.2420140846e-1f);
    p = fmaf(p, x, 1.4249322787e-1f);
    p = fmaf(p, x, -1.6668057665e-1f);
    p = fmaf(p, x, 2.0000714765e-1f);
    p = fmaf(p, x, -2.4999993993e-1f);
    p = fmaf(p, x, 3.3333331174e-1f);
    float y = (p * x) * z;
    float fe = (float)e;
    y = fmaf(fe, -2.12194440e-4f, y);
    y = fmaf(-0.5f, z, y);
    float r = x + y;
    r = fmaf(fe, 0.693359375f, r);
    return r;
}

static inline float om_powf(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : INFINITY;
    if (x < 0.0f) return (x - x) / (x - x);
    return om_expf(y * om_logf(x));
}

#endif /* NOOR_ORACLE_OMATH_H */
