/*
 * mmc_fp32math.h -- deterministic single-precision elementary functions.
 *
 * The photon kernel needs exp, log, sin/cos and (for isotropic media) acos in
 * FP32.  Vendor libraries (glibc on the CPU, the device math library on the
 * GPU) round these differently in the last bit, which is enough to make two
 * runs of the same photon drift apart.  These implementations use only IEEE
 * add, multiply, divide and sqrt plus integer bit operations, so a C compiler
 * and the HIP compiler produce bit-identical results when both compile with
 * -ffp-contract=off.  The CPU oracle can therefore replay any photon of the
 * GPU exactly.
 *
 * Method (textbook range reduction + Taylor polynomials):
 *   exp:  x = k ln2 + r (two-part Cody-Waite), e^r by its Taylor series to r^7
 *   log:  x = 2^k m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(f/(2+f))
 *   sincos: x = j pi/2 + r (three-part Cody-Waite), Taylor series to r^10
 *   acos: |x|<=1/2 via pi/2 - asin(x), else 2 asin(sqrt((1-|x|)/2))
 * Accuracy against double precision is checked in tests/test_fp32math.py.
 *
 * mmc_twopi_u reproduces (float)(2*M_PI*(double)u) for the random numbers
 * u = m*2^-23 that xorshift128+ produces, using integer arithmetic only (the
 * reference forms the azimuth in double, src/mmc_mesh.c:1666; the kernel
 * keeps all per-photon arithmetic in 32-bit registers).
 */
#ifndef MMC_FP32MATH_H
#define MMC_FP32MATH_H

#include <stdint.h>
#include <string.h>

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
    float r = x - kf * MMC_LN2_HI;
    r = r - kf * MMC_LN2_LO;
    float p = 1.98412701e-4f;               /* 1/7! */
    p = 1.38888892e-3f + r * p;             /* 1/6! */
    p = 8.33333377e-3f + r * p;             /* 1/5! */
    p = 4.16666679e-2f + r * p;             /* 1/4! */
    p = 1.66666672e-1f + r * p;             /* 1/3! */
    p = 0.5f + r * p;
    p = 1.0f + r * p;
    p = 1.0f + r * p;
    int k1 = k / 2, k2 = k - k1;            /* 2^k as two normal factors */
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
    float R = 0.222222224f;                 /* 2/9 */
    R = 0.285714298f + z * R;               /* 2/7 */
    R = 0.400000006f + z * R;               /* 2/5 */
    R = 0.666666687f + z * R;               /* 2/3 */
    R = z * R;
    float hfsq = 0.5f * f * f;
    float kf = (float)k;
    return kf * MMC_LN2_HI + ((f - (hfsq - s * (hfsq + R))) + kf * MMC_LN2_LO);
}

#define MMC_PIO2_1 1.5703125f                 /* 8 significant bits */
#define MMC_PIO2_2 4.83751296997070312e-4f    /* 12 significant bits */
#define MMC_PIO2_3 7.54979012640433210e-8f

/* sin and cos of |x| < 2^15 */
MMC_MATHFN void mmc_sincosf(float x, float* sp, float* cp) {
    float j = mmc_rint(x * 0.636619747f);
    int q = (int)j;
    float r = x - j * MMC_PIO2_1;
    r = r - j * MMC_PIO2_2;
    r = r - j * MMC_PIO2_3;
    float z = r * r;
    float ps = 2.75573188e-6f;              /* 1/9! */
    ps = -1.98412701e-4f + z * ps;          /* -1/7! */
    ps = 8.33333377e-3f + z * ps;           /* 1/5! */
    ps = -1.66666672e-1f + z * ps;          /* -1/3! */
    float s = r + (r * z) * ps;
    float pc = -2.75573188e-7f;             /* -1/10! */
    pc = 2.48015876e-5f + z * pc;           /* 1/8! */
    pc = -1.38888892e-3f + z * pc;          /* -1/6! */
    pc = 4.16666679e-2f + z * pc;           /* 1/4! */
    pc = -0.5f + z * pc;
    float c = 1.0f + z * pc;
    switch (q & 3) {
        case 0: *sp = s; *cp = c; break;
        case 1: *sp = c; *cp = -s; break;
        case 2: *sp = -s; *cp = -c; break;
        default: *sp = -c; *cp = s; break;
    }
}

/* asin for |y| <= 1/2: Taylor series sum_n (2n)!/(4^n (n!)^2 (2n+1)) y^(2n+1) */
MMC_MATHFN float mmc_asin_half(float y) {
    float z = y * y;
    float p = 6.44721044e-3f;
    p = 7.31252599e-3f + z * p;
    p = 8.39033537e-3f + z * p;
    p = 9.76160914e-3f + z * p;
    p = 1.15518011e-2f + z * p;
    p = 1.39648439e-2f + z * p;
    p = 1.73527643e-2f + z * p;
    p = 2.23721586e-2f + z * p;
    p = 3.03819440e-2f + z * p;
    p = 4.46428582e-2f + z * p;
    p = 7.50000030e-2f + z * p;
    p = 1.66666672e-1f + z * p;
    return y + (y * z) * p;
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

/* (float)(2*pi*(double)u) for u = m * 2^-23, 0 <= m < 2^23: the double
 * product rounded to 53 bits, then to 24 bits, both to nearest even. */
MMC_MATHFN int mmc_bitlen64(uint64_t t) { return t ? 64 - __builtin_clzll(t) : 0; }

MMC_MATHFN float mmc_twopi_u(uint32_t m) {
    if (m == 0) return 0.f;
    const uint64_t M = 0x1921FB54442D18ull;           /* 2*pi = M * 2^-50 (53 bits) */
    uint64_t lo = (M & 0xFFFFFFFFull) * m;
    uint64_t hi = (M >> 32) * m;                      /* product = hi*2^32 + lo < 2^76 */
    uint64_t mid = (lo >> 32) + (hi & 0xFFFFFFFFull);
    uint64_t top = (hi >> 32) + (mid >> 32);          /* bits 64..75 */
    uint64_t low = (mid << 32) | (lo & 0xFFFFFFFFull); /* bits 0..63 */
    int L = top ? 64 + mmc_bitlen64(top) : mmc_bitlen64(low);
    int e = -73;                                      /* value = (top:low) * 2^e */
    uint64_t v = low;
    if (L > 53) {                                     /* round to 53 bits, sh <= 23 */
        int sh = L - 53;
        uint64_t q = (low >> sh) | (top << (64 - sh));
        uint64_t r = low & ((1ull << sh) - 1), half = 1ull << (sh - 1);
        if (r > half || (r == half && (q & 1))) q++;
        v = q;
        e += sh;
    }
    int L2 = mmc_bitlen64(v);
    if (L2 > 24) {                                    /* round to 24 bits */
        int sh = L2 - 24;
        uint64_t q = v >> sh, r = v & ((1ull << sh) - 1), half = 1ull << (sh - 1);
        if (r > half || (r == half && (q & 1))) q++;
        v = q;
        e += sh;
    }
    int L3 = mmc_bitlen64(v);
    int expo = e + L3 - 1;                            /* a normal float for m >= 1 */
    uint32_t mant = (uint32_t)((v << (24 - L3)) & 0x7FFFFFull);
    return mmc_bitsf(((uint32_t)(expo + 127) << 23) | mant);
}

#endif
