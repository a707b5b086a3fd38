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

/* ---- double-precision evaluations of the generator-driven draws ----------
 * The host C library computes logf, sinf/cosf and acosf of a float argument
 * by evaluating in double and rounding to float; its results are within one
 * ulp of the exact value but not always correctly rounded.  The functions
 * below evaluate the same quantities in double to about 2^-50 relative error
 * (Taylor polynomials, Horner with explicit fused multiply-adds; IEEE
 * add/multiply/divide only), so their rounding to float is the C library's
 * result except near a rounding midpoint, where a table built from the C
 * library decides (mmc_host_libm_table).  Host and GPU compute bit-identical
 * doubles here, which is what lets a host-built table steer the GPU.
 *
 * The coefficients are read from a table K (MMC_DCOEF_LIST) rather than
 * written as literals: on the GPU, literal double constants are loop
 * invariants the compiler keeps in registers for the whole photon loop, while
 * table entries are fetched by scalar loads where they are used. */
#define MMC_DFMA(a, b, c) __builtin_fma((a), (b), (c))

/* K[0..10]   1/23, 1/21, ..., 1/3          (2 atanh series of log)
 * K[11..18]  1/17!, -1/15!, ..., -1/3!     (sin)
 * K[19..26]  -1/18!, 1/16!, ..., 1/4!      (cos)
 * K[27..48]  asin series, n = 22 .. 1: (2n)!/(4^n (n!)^2 (2n+1))
 * K[49..50]  ln 2 = K[49] + K[50]  (K[49]: 42 significant bits)
 * K[51..52]  pi/2 = K[51] + K[52];  K[53] = 2/pi
 * K[54..55]  pi = K[54] + K[55] */
#define MMC_NDCOEF 56
#define MMC_DCOEF_LIST                                                                      \
    0x1.642c8590b2164p-5, 0x1.8618618618618p-5, 0x1.af286bca1af28p-5, 0x1.e1e1e1e1e1e1ep-5, \
    0x1.1111111111111p-4, 0x1.3b13b13b13b14p-4, 0x1.745d1745d1746p-4, 0x1.c71c71c71c71cp-4, \
    0x1.2492492492492p-3, 0x1.999999999999ap-3, 0x1.5555555555555p-2,                       \
    0x1.952c77030ad4ap-49, -0x1.ae7f3e733b81fp-41, 0x1.6124613a86d09p-33,                  \
    -0x1.ae64567f544e4p-26, 0x1.71de3a556c734p-19, -0x1.a01a01a01a01ap-13,                 \
    0x1.1111111111111p-7, -0x1.5555555555555p-3,                                           \
    -0x1.6827863b97d97p-53, 0x1.ae7f3e733b81fp-45, -0x1.93974a8c07c9dp-37,                 \
    0x1.1eed8eff8d898p-29, -0x1.27e4fb7789f5cp-22, 0x1.a01a01a01a01ap-16,                  \
    -0x1.6c16c16c16c17p-10, 0x1.5555555555555p-5,                                          \
    0x1.5c5f56efaaaabp-9, 0x1.750de64d7d05fp-9, 0x1.90cb77f60c7cep-9, 0x1.b026f57b13b14p-9, \
    0x1.d3d2a8e0dd67dp-9, 0x1.fcaf8fb6db6dbp-9, 0x1.15ee9d45d1746p-8, 0x1.31683bdef7bdfp-8, \
    0x1.51ba308d3dcb1p-8, 0x1.782dda12f684cp-8, 0x1.a6863d70a3d71p-8, 0x1.df3bd37a6f4dfp-8, \
    0x1.12ef3cf3cf3cfp-7, 0x1.3fde50d79435ep-7, 0x1.7a87878787878p-7, 0x1.c99999999999ap-7, \
    0x1.1c4ec4ec4ec4fp-6, 0x1.6e8ba2e8ba2e9p-6, 0x1.f1c71c71c71c7p-6, 0x1.6db6db6db6db7p-5, \
    0x1.3333333333333p-4, 0x1.5555555555555p-3,                                            \
    0x1.62e42fefa3800p-1, 0x1.ef35793c76730p-45,                                           \
    0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54, 0x1.45f306dc9c883p-1,                     \
    0x1.921fb54442d18p+1, 0x1.1a62633145c07p-53

/* log(x) for a normal float x > 0: x = 2^k m, m in [sqrt(1/2), sqrt(2)),
 * s = (m - 1)/(m + 1), log m = 2 atanh(s) = 2 sum_n s^(2n+1)/(2n+1) to n = 11
 * (|s| <= 0.1716, so the first omitted term is below 2^-56 of the sum) */
MMC_MATHFN double mmc_log_dk(float x, const double* K) {
    const uint32_t ix = mmc_fbits(x);
    int k = (int)(ix >> 23) - 127;
    uint32_t mb = (ix & 0x007FFFFFu) | 0x3F800000u;
    if (mb > 0x3FB504F3u) { /* m > sqrt(2): halve it */
        mb -= 0x00800000u;
        k += 1;
    }
    const double f = (double)mmc_bitsf(mb) - 1.0;   /* exact */
    const double s = f / (2.0 + f);                  /* 2 + f exact */
    const double z = s * s;
    double p = K[0];
    for (int i = 1; i <= 10; i++) p = MMC_DFMA(z, p, K[i]);
    const double s2 = s + s;
    const double logm = MMC_DFMA(s2 * z, p, s2);
    const double kd = (double)k;
    return MMC_DFMA(kd, K[49], MMC_DFMA(kd, K[50], logm));
}

/* sin and cos of a float |a| < 2^20: a = j pi/2 + r, |r| <= pi/4 (pi/2 in
 * two parts; each fused multiply-add rounds once), Taylor series to r^17 and
 * r^18 */
MMC_MATHFN void mmc_sincos_dk(float a, double* sp, double* cp, const double* K) {
    const double x = (double)a;
    const double big = 0x1.8p52;
    const double j = (x * K[53] + big) - big;        /* nearest integer */
    const int q = (int)j;
    double r = MMC_DFMA(-j, K[51], x);
    r = MMC_DFMA(-j, K[52], r);
    const double z = r * r;
    double ps = K[11];
    for (int i = 12; i <= 18; i++) ps = MMC_DFMA(z, ps, K[i]);
    const double s = MMC_DFMA(r * z, ps, r);
    double pc = K[19];
    for (int i = 20; i <= 26; i++) pc = MMC_DFMA(z, pc, K[i]);
    pc = MMC_DFMA(z, pc, -0.5);
    const double c = MMC_DFMA(z, pc, 1.0);
    const double s2 = (q & 1) ? c : s, c2 = (q & 1) ? s : c;
    *sp = (q & 2) ? -s2 : s2;
    *cp = ((q + 1) & 2) ? -c2 : c2;
}

/* asin(y) for |y| <= 1/2: sum_n (2n)!/(4^n (n!)^2 (2n+1)) y^(2n+1) to n = 22 */
MMC_MATHFN double mmc_asin_half_dk(double y, const double* K) {
    const double z = y * y;
    double p = K[27];
    for (int i = 28; i <= 48; i++) p = MMC_DFMA(z, p, K[i]);
    return MMC_DFMA(y * z, p, y);
}

/* acos(x) for a float -1 <= x <= 1: pi/2 - asin(x) for |x| <= 1/2, else
 * 2 asin(sqrt((1 - |x|)/2)) (from pi for x < 0).  (1 - |x|)/2 is a float, its
 * square root is the correctly rounded float root refined by one Newton step
 * in double. */
MMC_MATHFN double mmc_acos_dk(float xf, const double* K) {
    const double x = (double)xf;
    if (xf <= 0.5f && xf >= -0.5f) return K[51] - (mmc_asin_half_dk(x, K) - K[52]);
    const float zf = (1.f - (xf > 0.f ? xf : -xf)) * 0.5f;      /* exact */
    const double z = (double)zf;
    const double s0 = (double)MMC_SQRTF(zf);
    const double t = (zf > 0.f) ? s0 + MMC_DFMA(-s0, s0, z) / (s0 + s0) : 0.0;
    const double a = 2.0 * mmc_asin_half_dk(t, K);
    return (xf > 0.f) ? a : K[54] - (a - K[55]);
}

#if !defined(__HIPCC__)
/* host C: the same evaluations with the coefficient table in static memory */
static const double mmc_dcoef_host[MMC_NDCOEF] = {MMC_DCOEF_LIST};
static inline double mmc_log_d(float x) { return mmc_log_dk(x, mmc_dcoef_host); }
static inline void mmc_sincos_d(float a, double* s, double* c) { mmc_sincos_dk(a, s, c, mmc_dcoef_host); }
static inline double mmc_acos_d(float x) { return mmc_acos_dk(x, mmc_dcoef_host); }
#endif

/* the float next to f, above it (up) or below it */
MMC_MATHFN float mmc_fstep(float f, int up) {
    const uint32_t b = mmc_fbits(f);
    if (f == 0.f) return mmc_bitsf(up ? 1u : 0x80000001u);
    return mmc_bitsf(((f > 0.f) == (up != 0)) ? b + 1u : b - 1u);
}
/* the float next to f in the direction of the double e (f = (float)e) */
MMC_MATHFN float mmc_fnext(float f, double e) { return mmc_fstep(f, e > (double)f); }

#endif
