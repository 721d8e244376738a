/*
 * mcx_fpmath.h - deterministic single-precision elementary functions.
 *
 * The photon-transport kernel can be built in two arithmetic modes:
 *
 *  - "fast": hardware transcendentals (v_exp_f32, v_log_f32, v_sin_f32, ...)
 *    and reciprocal-based division, the MI355X analogue of the reference's
 *    native_log/native_exp/native_divide (src/mcx_core.cl:158-173, built with
 *    -cl-mad-enable -DMCX_USE_NATIVE, src/mcx_host.cpp:857-858).
 *  - "parity": every transcendental is evaluated by the functions below, which
 *    use only IEEE-754 single-precision add, subtract, multiply, divide and
 *    square root (all correctly rounded on x86-64 SSE and on gfx950 with the
 *    correctly-rounded divide/sqrt lowering), integer bit manipulation and
 *    comparisons.  Compiled without floating-point contraction, the same input
 *    gives the same bits on the host CPU and on the GPU, so a photon history
 *    computed by the CPU oracle and by the gfx950 kernel agree bit for bit.
 *
 * Methods (all textbook or published, written from the mathematics):
 *  - exp:  Cody-Waite reduction x = k ln2 + r with a two-part ln2, degree-7
 *          Taylor polynomial for e^r on |r| <= ln2/2, exponent scaling by
 *          integer bit insertion.
 *  - log:  x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 *          s = (m-1)/(m+1), odd series to s^11.
 *  - sin/cos: Cody-Waite reduction by pi/2 with a three-part constant,
 *          Taylor polynomials on |r| <= pi/4, quadrant selection.
 *  - acos: Cephes single-precision asin kernel (S. L. Moshier, Cephes Math
 *          Library, asinf.c: degree-5 polynomial in x^2 on |x| <= 1/2) with
 *          the half-angle identity acos(x) = 2 asin(sqrt((1-x)/2)) for x > 1/2
 *          and acos(x) = pi - 2 asin(sqrt((1+x)/2)) for x < -1/2.
 *
 * Accuracy against double-precision libm is checked in tests/test_oracle_pins.py.
 */
#ifndef MCX_FPMATH_H
#define MCX_FPMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define MCX_HD __host__ __device__ __forceinline__
#else
#define MCX_HD static inline
#endif

typedef union {
    float f;
    uint32_t u;
    int32_t i;
} mcx_fbits;

MCX_HD float pm_asfloat(uint32_t u) {
    mcx_fbits b;
    b.u = u;
    return b.f;
}

MCX_HD uint32_t pm_asuint(float f) {
    mcx_fbits b;
    b.f = f;
    return b.u;
}

MCX_HD float pm_fabs(float x) {
    return pm_asfloat(pm_asuint(x) & 0x7FFFFFFFu);
}

/* round to nearest integer (ties away from zero is irrelevant: only used on
   reduced arguments); returns an int, exact for |x| < 2^30 */
MCX_HD int pm_rint_i(float x) {
    return (int)(x >= 0.f ? x + 0.5f : x - 0.5f);
}

/* 2^k for integer k in [-126, 127] as a float */
MCX_HD float pm_pow2i(int k) {
    return pm_asfloat((uint32_t)(k + 127) << 23);
}

MCX_HD float pm_exp(float x) {
    const float LOG2E = 1.44269504088896341f;
    const float LN2_HI = 0.693359375f;        /* 355/512, exact */
    const float LN2_LO = -2.12194440054690583e-4f;

    if (!(x == x)) {
        return x;
    }

    if (x > 88.7228317f) {
        return pm_asfloat(0x7F800000u);
    }

    if (x < -103.972084f) {
        return 0.f;
    }

    int k = pm_rint_i(x * LOG2E);
    float kf = (float)k;
    float r = (x - kf * LN2_HI) - kf * LN2_LO;
    /* e^r, Horner form of the Taylor series to r^7 */
    float p = 1.984126984e-4f;                /* 1/5040 */
    p = p * r + 1.388888889e-3f;              /* 1/720 */
    p = p * r + 8.333333333e-3f;              /* 1/120 */
    p = p * r + 4.166666667e-2f;              /* 1/24 */
    p = p * r + 1.666666667e-1f;              /* 1/6 */
    p = p * r + 0.5f;
    p = p * r + 1.f;
    p = p * r + 1.f;

    /* scale by 2^k, splitting the scale when the result is subnormal */
    if (k < -125) {
        return (p * pm_pow2i(k + 64)) * pm_pow2i(-64);
    }

    if (k > 127) {
        return (p * pm_pow2i(k - 1)) * 2.f;
    }

    return p * pm_pow2i(k);
}

MCX_HD float pm_log(float x) {
    const float LN2_HI = 0.693359375f;
    const float LN2_LO = -2.12194440054690583e-4f;

    if (!(x > 0.f)) {
        return (x == 0.f) ? pm_asfloat(0xFF800000u) : pm_asfloat(0x7FC00000u);
    }

    if (x == pm_asfloat(0x7F800000u)) {
        return x;
    }

    int e = 0;

    if (x < 1.17549435e-38f) {   /* subnormal: bring into the normal range */
        x *= 8388608.f;          /* 2^23 */
        e = -23;
    }

    uint32_t ix = pm_asuint(x);
    e += (int)(ix >> 23) - 127;
    /* mantissa in [1,2) */
    float m = pm_asfloat((ix & 0x007FFFFFu) | 0x3F800000u);

    if (m > 1.41421356f) {
        m *= 0.5f;
        e += 1;
    }

    float f = m - 1.f;
    float s = f / (2.f + f);
    float z = s * s;
    float q = 9.090909091e-2f;                /* 1/11 */
    q = q * z + 1.111111111e-1f;              /* 1/9 */
    q = q * z + 1.428571429e-1f;              /* 1/7 */
    q = q * z + 2.e-1f;                       /* 1/5 */
    q = q * z + 3.333333333e-1f;              /* 1/3 */
    q = q * z;
    float s2 = s + s;
    float lnm = s2 + s2 * q;
    float ef = (float)e;
    return ef * LN2_HI + (ef * LN2_LO + lnm);
}

/* sin and cos of x by reduction modulo pi/2; accurate for |x| < 1e4 */
MCX_HD void pm_sincos(float x, float* s, float* c) {
    const float TWO_OVER_PI = 0.636619772367581343f;
    const float PIO2_1 = 1.5703125f;                  /* exact, 8 bits */
    const float PIO2_2 = 4.83751296997070312e-4f;     /* exact, 12 bits */
    const float PIO2_3 = 7.54978995489188216e-8f;

    int k = pm_rint_i(x * TWO_OVER_PI);
    float kf = (float)k;
    float r = ((x - kf * PIO2_1) - kf * PIO2_2) - kf * PIO2_3;
    float r2 = r * r;

    /* sin(r) = r - r^3/3! + ... + r^9/9! */
    float ps = 2.755731922e-6f;               /* 1/9! */
    ps = ps * r2 - 1.984126984e-4f;           /* 1/7! */
    ps = ps * r2 + 8.333333333e-3f;           /* 1/5! */
    ps = ps * r2 - 1.666666667e-1f;           /* 1/3! */
    float sr = r + (r * r2) * ps;

    /* cos(r) = 1 - r^2/2! + ... - r^10/10! */
    float pc = -2.755731922e-7f;              /* 1/10! */
    pc = pc * r2 + 2.480158730e-5f;           /* 1/8! */
    pc = pc * r2 - 1.388888889e-3f;           /* 1/6! */
    pc = pc * r2 + 4.166666667e-2f;           /* 1/4! */
    pc = pc * r2 - 0.5f;
    float cr = 1.f + r2 * pc;

    switch (k & 3) {
        case 0:
            *s = sr;
            *c = cr;
            break;

        case 1:
            *s = cr;
            *c = -sr;
            break;

        case 2:
            *s = -sr;
            *c = -cr;
            break;

        default:
            *s = -cr;
            *c = sr;
            break;
    }
}

MCX_HD float pm_sin(float x) {
    float s, c;
    pm_sincos(x, &s, &c);
    return s;
}

MCX_HD float pm_cos(float x) {
    float s, c;
    pm_sincos(x, &s, &c);
    return c;
}

/* Cephes asinf kernel, valid for |x| <= 0.5 */
MCX_HD float pm_asin_kernel(float x) {
    float z = x * x;
    float p = 4.2163199048e-2f;
    p = p * z + 2.4181311049e-2f;
    p = p * z + 4.5470025998e-2f;
    p = p * z + 7.4953002686e-2f;
    p = p * z + 1.6666752422e-1f;
    return (p * z) * x + x;
}

MCX_HD float pm_acos(float x) {
    const float PI = 3.14159265358979f;
    const float PIO2 = 1.57079632679490f;

    if (x > 0.5f) {
        float z = 0.5f * (1.f - x);
        return 2.f * pm_asin_kernel(__builtin_sqrtf(z));
    }

    if (x < -0.5f) {
        float z = 0.5f * (1.f + x);
        return PI - 2.f * pm_asin_kernel(__builtin_sqrtf(z));
    }

    return PIO2 - pm_asin_kernel(x);
}

MCX_HD float pm_sqrt(float x) {
    return __builtin_sqrtf(x);
}

MCX_HD float pm_rsqrt(float x) {
    return 1.f / __builtin_sqrtf(x);
}

#endif
