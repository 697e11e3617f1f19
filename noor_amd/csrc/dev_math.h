// Portable single-precision transcendentals for the gfx950 kernels.
//
// The render path needs sin, cos, atan2, acos, exp, log and pow.  The vendor
// device library and the host C library round these differently by an ulp or
// two, which is enough to flip Russian-roulette and lobe-selection branches.
// These implementations use only IEEE +, -, *, /, sqrt, floor and fma, so
// they round identically on the CPU and the GPU:
//   - Cody-Waite reduction by pi/4 (three-part constant) and the
//     single-precision minimax polynomials of S. L. Moshier's Cephes library
//     for sin, cos, atan, asin, exp and log;
//   - acos through the half-angle identity, atan2 from atan, pow = exp(y log x).
// Valid for |x| <= 8192 in sin / cos, which covers every angle the renderer
// forms (at most a few turns).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace noor {

#define NM_PI 3.14159265358979323846f
#define NM_PI_2 1.57079632679489661923f
#define NM_PI_4 0.785398163397448309616f

__device__ __forceinline__ float nm_ldexp(float x, int n) {
    const int n1 = n / 2;
    const int n2 = n - n1;
    const float s1 = __uint_as_float((uint32_t)(n1 + 127) << 23);
    const float s2 = __uint_as_float((uint32_t)(n2 + 127) << 23);
    return (x * s1) * s2;
}

__device__ __forceinline__ float nm_sin_poly(float x) {
    const float z = x * x;
    float p = __builtin_fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f);
    p = __builtin_fmaf(p, z, -1.6666654611e-1f);
    return __builtin_fmaf(p * z, x, x);
}

__device__ __forceinline__ float nm_cos_poly(float x) {
    const float z = x * x;
    float p = __builtin_fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f);
    p = __builtin_fmaf(p, z, 4.166664568298827e-2f);
    return __builtin_fmaf(p * z, z, __builtin_fmaf(-0.5f, z, 1.0f));
}

__device__ __forceinline__ int nm_reduce(float x, float* r) {
    const float ax = fabsf(x);
    int j = (int)(ax * 1.27323954473516f);
    j += (j & 1);
    const float y = (float)j;
    float t = __builtin_fmaf(-y, 0.78515625f, ax);
    t = __builtin_fmaf(-y, 2.4187564849853515625e-4f, t);
    t = __builtin_fmaf(-y, 3.77489497744594108e-8f, t);
    *r = t;
    return j;
}

__device__ __forceinline__ float nm_sinf(float x) {
    if (!(fabsf(x) <= 3.0e38f)) return x - x;
    float r;
    const int j = nm_reduce(x, &r);
    const int q = (j >> 1) & 3;
    float v = (q & 1) ? nm_cos_poly(r) : nm_sin_poly(r);
    if (q & 2) v = -v;
    return (x < 0.0f) ? -v : v;
}

__device__ __forceinline__ float nm_cosf(float x) {
    if (!(fabsf(x) <= 3.0e38f)) return x - x;
    float r;
    const int j = nm_reduce(x, &r);
    const int q = (j >> 1) & 3;
    float v = (q & 1) ? nm_sin_poly(r) : nm_cos_poly(r);
    if (q == 1 || q == 2) v = -v;
    return v;
}

__device__ __forceinline__ float nm_atanf(float x) {
    float sign = 1.0f;
    if (x < 0.0f) {
        sign = -1.0f;
        x = -x;
    }
    float y;
    if (x > 2.414213562373095f) {
        y = NM_PI_2;
        x = -1.0f / x;
    } else if (x > 0.4142135623730950f) {
        y = NM_PI_4;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    const float z = x * x;
    float p = __builtin_fmaf(8.05374449538e-2f, z, -1.38776856032e-1f);
    p = __builtin_fmaf(p, z, 1.99777106478e-1f);
    p = __builtin_fmaf(p, z, -3.33329491539e-1f);
    y = y + __builtin_fmaf(p * z, x, x);
    return sign * y;
}

__device__ __forceinline__ float nm_atan2f(float y, float x) {
    if (x != x || y != y) return x + y;
    if (x == 0.0f) {
        if (y > 0.0f) return NM_PI_2;
        if (y < 0.0f) return -NM_PI_2;
        return 0.0f;
    }
    if (y == 0.0f) return (x > 0.0f) ? 0.0f : NM_PI;
    float w = 0.0f;
    if (x < 0.0f) w = (y < 0.0f) ? -NM_PI : NM_PI;
    return w + nm_atanf(y / x);
}

__device__ __forceinline__ float nm_asinf(float x) {
    const float a = fabsf(x);
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
    float p = __builtin_fmaf(4.2163199048e-2f, z, 2.4181311049e-2f);
    p = __builtin_fmaf(p, z, 4.5470025998e-2f);
    p = __builtin_fmaf(p, z, 7.4953002686e-2f);
    p = __builtin_fmaf(p, z, 1.6666752422e-1f);
    float r = __builtin_fmaf(p * z, t, t);
    if (flag) r = NM_PI_2 - (r + r);
    return (x < 0.0f) ? -r : r;
}

__device__ __forceinline__ float nm_acosf(float x) {
    if (x != x || fabsf(x) > 1.0f) return (x - x) / (x - x);
    if (x < -0.5f) return NM_PI - 2.0f * nm_asinf(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * nm_asinf(sqrtf(0.5f * (1.0f - x)));
    return NM_PI_2 - nm_asinf(x);
}

__device__ __forceinline__ float nm_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283905206835f) return __builtin_huge_valf();
    if (x < -103.278929903431851103f) return 0.0f;
    const float z = floorf(__builtin_fmaf(1.44269504088896341f, x, 0.5f));
    const int n = (int)z;
    x = __builtin_fmaf(-z, 0.693359375f, x);
    x = __builtin_fmaf(-z, -2.12194440e-4f, x);
    float p = __builtin_fmaf(1.9875691500e-4f, x, 1.3981999507e-3f);
    p = __builtin_fmaf(p, x, 8.3334519073e-3f);
    p = __builtin_fmaf(p, x, 4.1665795894e-2f);
    p = __builtin_fmaf(p, x, 1.6666665459e-1f);
    p = __builtin_fmaf(p, x, 5.0000001201e-1f);
    const float xx = x * x;
    const float r = __builtin_fmaf(p, xx, x) + 1.0f;
    return nm_ldexp(r, n);
}

__device__ __forceinline__ float nm_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return (x - x) / (x - x);
    if (x == 0.0f) return -__builtin_huge_valf();
    if (x == __builtin_huge_valf()) return x;
    int e = 0;
    uint32_t u = __float_as_uint(x);
    if ((u >> 23) == 0) {
        x = x * 33554432.0f;
        u = __float_as_uint(x);
        e = -25;
    }
    e += (int)(u >> 23) - 126;
    x = __uint_as_float((u & 0x007FFFFFu) | 0x3F000000u);
    if (x < 0.707106781186547524f) {
        e -= 1;
        x = x + x - 1.0f;
    } else {
        x = x - 1.0f;
    }
    const float z = x * x;
    float p = __builtin_fmaf(7.0376836292e-2f, x, -1.1514610310e-1f);
    p = __builtin_fmaf(p, x, 1.1676998740e-1f);
    p = __builtin_fmaf(p, x, -1.2420140846e-1f);
    p = __builtin_fmaf(p, x, 1.4249322787e-1f);
    p = __builtin_fmaf(p, x, -1.6668057665e-1f);
    p = __builtin_fmaf(p, x, 2.0000714765e-1f);
    p = __builtin_fmaf(p, x, -2.4999993993e-1f);
    p = __builtin_fmaf(p, x, 3.3333331174e-1f);
    float y = (p * x) * z;
    const float fe = (float)e;
    y = __builtin_fmaf(fe, -2.12194440e-4f, y);
    y = __builtin_fmaf(-0.5f, z, y);
    float r = x + y;
    r = __builtin_fmaf(fe, 0.693359375f, r);
    return r;
}

__device__ __forceinline__ float nm_powf(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : __builtin_huge_valf();
    if (x < 0.0f) return (x - x) / (x - x);
    return nm_expf(y * nm_logf(x));
}

}  // namespace noor
