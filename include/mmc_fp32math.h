/*
 * mmc_fp32math.h -- deterministic single-precision elementary functions.
 *
 * The photon kernel needs exp, log, sin/cos and (for isotropic media) acos in
 * FP32.  Vendor libraries (glibc on the CPU, the device math library on the
 * GPU) round these differently in the last bit, which is enough to make two
 * runs of the same photon drift apart.  These implementations use only IEEE
 * add, multiply, fused multiply-add, divide and sqrt plus integer bit
 * operations, so a C compiler and the HIP compiler produce bit-identical
 * results when both compile with -ffp-contract=off (the fused operations are
 * written out explicitly).  The CPU oracle can therefore replay any photon of
 * the GPU exactly.
 *
 * Method (textbook range reduction + Taylor polynomials, Horner with FMA):
 *   exp:  x = k ln2 + r (two-part Cody-Waite), e^r by its Taylor series to r^7
 *   log:  x = 2^k m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(f/(2+f))
 *   sincos: x = j pi/2 + r (three-part Cody-Waite), Taylor series to r^10
 *   acos: |x|<=1/2 via pi/2 - asin(x), else 2 asin(sqrt((1-|x|)/2))
 * Accuracy against double precision is checked in tests/test_fp32math.py.
 *
 * mmc_twopi is the one double-precision operation of the photon loop: the
 * reference forms the scattering azimuth as a double product
 * (src/mmc_mesh.c:1666), and gfx950 issues a v_mul_f64 at the FP32 rate.
 */
#ifndef MMC_FP32MATH_H
#define MMC_FP32MATH_H

#include <stdint.h>
#include <string.h>

/* fused multiply-add: one rounding, identical in C (fmaf) and on the GPU */
#define MMC_FMA(a, b, c) __builtin_fmaf((a), (b), (c))

#if defined(__HIPCC__)
#define MMC_MATHFN __host__ __device__ __forceinline__
#define MMC_SQRTF(x) sqrtf(x)
#else
#define MMC_MATHFN static inline
#define MMC_SQRTF(x) __builtin_sqrtf(x)
#endif

MMC_MATHFN uint32_t mmc_fbits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
MMC_MATHFN float mmc_bitsf(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
/* round to nearest even for |x| < 2^22 without relying on the rounding mode */
MMC_MATHFN float mmc_rint(float x) {
    const float big = 12582912.0f; /* 1.5 * 2^23 */
    return (x + big) - big;
}
MMC_MATHFN float mmc_pow2i(int k) { /* 2^k for k in [-126, 127] */
    return mmc_bitsf((uint32_t)(k + 127) << 23);
}

#define MMC_LN2_HI 0.693359375f             /* 9 significant bits */
#define MMC_LN2_LO -2.12194441701285541e-4f /* ln2 - MMC_LN2_HI */

MMC_MATHFN float mmc_expf(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return mmc_bitsf(0x7F800000u);
    if (x < -103.972084f) return 0.f;
    float kf = mmc_rint(x * 1.44269502f);
    int k = (int)kf;
    float r = MMC_FMA(-kf, MMC_LN2_HI, x);   /* exact: kf*LN2_HI has <= 17 bits */
    r = MMC_FMA(-kf, MMC_LN2_LO, r);
    float p = 1.98412701e-4f;                /* 1/7! */
    p = MMC_FMA(r, p, 1.38888892e-3f);       /* 1/6! */
    p = MMC_FMA(r, p, 8.33333377e-3f);       /* 1/5! */
    p = MMC_FMA(r, p, 4.16666679e-2f);       /* 1/4! */
    p = MMC_FMA(r, p, 1.66666672e-1f);       /* 1/3! */
    p = MMC_FMA(r, p, 0.5f);
    p = MMC_FMA(r, p, 1.0f);
    p = MMC_FMA(r, p, 1.0f);
    int k1 = k >> 1, k2 = k - k1;            /* 2^k as two normal factors */
    return (p * mmc_pow2i(k1)) * mmc_pow2i(k2);
}

MMC_MATHFN float mmc_logf(float x) {
    if (x != x) return x;
    if (x <= 0.f) return x == 0.f ? mmc_bitsf(0xFF800000u) : mmc_bitsf(0x7FC00000u);
    uint32_t ix = mmc_fbits(x);
    if (ix >= 0x7F800000u) return x;
    int k = 0;
    if (ix < 0x00800000u) { /* subnormal */
        x = x * 33554432.0f; /* 2^25 */
        ix = mmc_fbits(x);
        k = -25;
    }
    k += (int)(ix >> 23) - 127;
    float m = mmc_bitsf((ix & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421354f) {
        m = m * 0.5f;
        k += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = 0.222222224f;                  /* 2/9 */
    R = MMC_FMA(z, R, 0.285714298f);         /* 2/7 */
    R = MMC_FMA(z, R, 0.400000006f);         /* 2/5 */
    R = MMC_FMA(z, R, 0.666666687f);         /* 2/3 */
    R = z * R;
    float hfsq = 0.5f * f * f;
    float kf = (float)k;
    float t = MMC_FMA(-s, hfsq + R, hfsq);   /* hfsq - s*(hfsq+R) */
    return MMC_FMA(kf, MMC_LN2_HI, MMC_FMA(kf, MMC_LN2_LO, f - t));
}

#define MMC_PIO2_1 1.5703125f                 /* 8 significant bits */
#define MMC_PIO2_2 4.83751296997070312e-4f    /* 12 significant bits */
#define MMC_PIO2_3 7.54979012640433210e-8f

/* sin and cos of |x| < 2^15 */
MMC_MATHFN void mmc_sincosf(float x, float* sp, float* cp) {
    float j = mmc_rint(x * 0.636619747f);
    int q = (int)j;
    float r = MMC_FMA(-j, MMC_PIO2_1, x);
    r = MMC_FMA(-j, MMC_PIO2_2, r);
    r = MMC_FMA(-j, MMC_PIO2_3, r);
    float z = r * r;
    float ps = 2.75573188e-6f;               /* 1/9! */
    ps = MMC_FMA(z, ps, -1.98412701e-4f);    /* -1/7! */
    ps = MMC_FMA(z, ps, 8.33333377e-3f);     /* 1/5! */
    ps = MMC_FMA(z, ps, -1.66666672e-1f);    /* -1/3! */
    float s = MMC_FMA(r * z, ps, r);
    float pc = -2.75573188e-7f;              /* -1/10! */
    pc = MMC_FMA(z, pc, 2.48015876e-5f);     /* 1/8! */
    pc = MMC_FMA(z, pc, -1.38888892e-3f);    /* -1/6! */
    pc = MMC_FMA(z, pc, 4.16666679e-2f);     /* 1/4! */
    pc = MMC_FMA(z, pc, -0.5f);
    float c = MMC_FMA(z, pc, 1.0f);
    float s2 = (q & 1) ? c : s, c2 = (q & 1) ? s : c;
    *sp = (q & 2) ? -s2 : s2;
    *cp = ((q + 1) & 2) ? -c2 : c2;
}

/* asin for |y| <= 1/2: Taylor series sum_n (2n)!/(4^n (n!)^2 (2n+1)) y^(2n+1) */
MMC_MATHFN float mmc_asin_half(float y) {
    float z = y * y;
    float p = 6.44721044e-3f;
    p = MMC_FMA(z, p, 7.31252599e-3f);
    p = MMC_FMA(z, p, 8.39033537e-3f);
    p = MMC_FMA(z, p, 9.76160914e-3f);
    p = MMC_FMA(z, p, 1.15518011e-2f);
    p = MMC_FMA(z, p, 1.39648439e-2f);
    p = MMC_FMA(z, p, 1.73527643e-2f);
    p = MMC_FMA(z, p, 2.23721586e-2f);
    p = MMC_FMA(z, p, 3.03819440e-2f);
    p = MMC_FMA(z, p, 4.46428582e-2f);
    p = MMC_FMA(z, p, 7.50000030e-2f);
    p = MMC_FMA(z, p, 1.66666672e-1f);
    return MMC_FMA(y * z, p, y);
}

MMC_MATHFN float mmc_acosf(float x) {
    const float pio2 = 1.57079637f, pi = 3.14159274f;
    if (x != x || x > 1.f || x < -1.f) return mmc_bitsf(0x7FC00000u);
    if (x <= 0.5f && x >= -0.5f) return pio2 - mmc_asin_half(x);
    if (x > 0.f) {
        float t = MMC_SQRTF((1.f - x) * 0.5f);
        return 2.f * mmc_asin_half(t);
    }
    float t = MMC_SQRTF((1.f + x) * 0.5f);
    return pi - 2.f * mmc_asin_half(t);
}

/* (float)(2*pi*(double)u): the reference forms the azimuth of a scattering
 * event in double precision (src/mmc_mesh.c:1666).  u is a float, so (double)u
 * is exact; the product is rounded to 53 bits and then to 24, both to nearest
 * even, on the CPU and the GPU alike. */
MMC_MATHFN float mmc_twopi(float u) { return (float)(6.283185307179586 * (double)u); }
/* (float)(pi*(double)u), the arcsine source's zenith angle
 * (src/mmc_raytrace.c:2478). */
MMC_MATHFN float mmc_pi(float u) { return (float)(3.141592653589793 * (double)u); }

#endif
