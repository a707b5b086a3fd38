/*
 * mmc_libm.c -- agreement tables between the GPU's draws and the host C
 * library's logf / sincosf / acosf.
 *
 * The reference CPU tracer draws every scattering length, azimuth and
 * isotropic zenith through the C library (src/mmc_rand_common.h:50-51,
 * src/mmc_mesh.c:1662-1690, src/mmc_raytrace.c:2434-2480), and a one-ulp
 * difference in any of them moves the photon for the rest of its life.  The
 * C library's float functions are within one ulp of the exact value but not
 * correctly rounded, so no independent implementation reproduces them by
 * construction.  Each of these calls sees only a function of one generator
 * output, u = k / 2^23 with k the 23 mantissa bits of xorshift128+
 * (src/mmc_rand_xorshift128p.c:55-70), so the host can enumerate them all.
 *
 * The GPU evaluates each value in double (mmc_log_d, mmc_sincos_d,
 * mmc_acos_d of mmc_fp32math.h; the host computes the same doubles bit for
 * bit) and rounds it to float, f.  The C library returns either f or the
 * float o on the other side of the double, next to f.  A bit per k and value
 * says which (set: o).  Only values close to the rounding midpoint between f
 * and o can be o, so each region of 2^MMC_LIBM_REGSHIFT consecutive k keeps a
 * threshold D per value: the largest distance |e - (f + o)/2|, in units of
 * |o - f|, over the region's set bits (rounded up to float; -1 when the
 * region has none).  The GPU reads the bit only when its double lies within D
 * of the midpoint -- a few percent of the draws -- and uses f otherwise.
 *
 * Tables (k = 0 .. 2^23-1), bits per k:
 *   MMC_FIX_LOG  logf(u + 1e-6f)                         1: log
 *   MMC_FIX_AZ   sincosf((float)(2 pi u))                2: sin, cos
 *   MMC_FIX_ISO  t = acosf(2u - 1); sincosf(t)           4: acos, sin, cos (of the
 *                                                           C library's t), unused
 *   MMC_FIX_PI   sincosf((float)(pi u))                  2: sin, cos
 *
 * Compiled with -mfma: the explicit fused multiply-adds of mmc_fp32math.h
 * become single instructions (the library's __builtin_fma would otherwise
 * call the C library's fma, which rounds identically but slowly).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "mmc_fp32math.h"
#include "mmc_hip_host.h"

#define NK (1u << 23)
#define NTHREAD 8

/* The argument of each table at generator output k, exactly as the kernel
 * forms it. */
static float u_of(uint32_t k) { return mmc_bitsf(0x3F800000u | k) - 1.0f; }

int32_t mmc_host_libm_nval(int32_t which) {
    switch (which) {
    case MMC_FIX_LOG: return 1;
    case MMC_FIX_AZ:
    case MMC_FIX_PI: return 2;
    case MMC_FIX_ISO: return 3;
    default: return -1;
    }
}

int64_t mmc_host_libm_table_bytes(int32_t which) {
    switch (which) {
    case MMC_FIX_LOG: return MMC_FIXB_LOG;
    case MMC_FIX_AZ: return MMC_FIXB_AZ;
    case MMC_FIX_PI: return MMC_FIXB_PI;
    case MMC_FIX_ISO: return MMC_FIXB_ISO;
    default: return -1;
    }
}

typedef struct {
    int32_t which;
    uint32_t k0, k1;
    uint8_t* bits;
    double dmax[(NK >> MMC_LIBM_REGSHIFT) / NTHREAD + 1][3]; /* per region of this job, per value */
    int64_t bad;
} fixjob;

/* records value v of generator output k: the C library's result g against
 * the double e; returns the bit (1: g is the far neighbour) */
static uint32_t classify(fixjob* j, uint32_t k, int v, double e, float g) {
    const float f = (float)e;
    if (mmc_fbits(g) == mmc_fbits(f)) return 0u;
    const float o = mmc_fnext(f, e);
    if (mmc_fbits(g) != mmc_fbits(o) || e == (double)f) {
        j->bad++;                         /* not reproducible by the rule */
        return 0u;
    }
    const double fd = f, od = o;
    const double q = fabs(e - 0.5 * (fd + od)) / fabs(od - fd);   /* exact: a power of two */
    const uint32_t r = (k >> MMC_LIBM_REGSHIFT) - (j->k0 >> MMC_LIBM_REGSHIFT);
    if (q > j->dmax[r][v]) j->dmax[r][v] = q;
    return 1u;
}

static void* fix_range(void* arg) {
    fixjob* j = (fixjob*)arg;
    for (uint32_t k = j->k0; k < j->k1; k++) {
        const float u = u_of(k);
        if (j->which == MMC_FIX_LOG) {
            const float x = u + 1e-6f;
            const uint32_t b = classify(j, k, 0, mmc_log_d(x), logf(x));
            j->bits[k >> 3] |= (uint8_t)(b << (k & 7));
        } else if (j->which == MMC_FIX_AZ || j->which == MMC_FIX_PI) {
            const float a = (j->which == MMC_FIX_AZ) ? mmc_twopi(u) : mmc_pi(u);
            float s, c;
            double es, ec;
            sincosf(a, &s, &c);
            mmc_sincos_d(a, &es, &ec);
            const uint32_t b = classify(j, k, 0, es, s) | (classify(j, k, 1, ec, c) << 1);
            j->bits[k >> 2] |= (uint8_t)(b << (2 * (k & 3)));
        } else { /* MMC_FIX_ISO: the C library's t feeds its sincosf */
            const float x = 2.f * u - 1.f;
            const float t = acosf(x);
            float s, c;
            double es, ec;
            sincosf(t, &s, &c);
            mmc_sincos_d(t, &es, &ec);
            const uint32_t b = classify(j, k, 0, mmc_acos_d(x), t) | (classify(j, k, 1, es, s) << 1) |
                               (classify(j, k, 2, ec, c) << 2);
            j->bits[k >> 1] |= (uint8_t)(b << (4 * (k & 1)));
        }
    }
    return NULL;
}

int32_t mmc_host_libm_table(int32_t which, uint8_t* bits, float* thr) {
    const int64_t nbytes = mmc_host_libm_table_bytes(which);
    const int nval = mmc_host_libm_nval(which);
    if (nbytes < 0 || !bits || !thr) return -1;
    memset(bits, 0, (size_t)nbytes);
    static fixjob jobs[NTHREAD];      /* large: kept off the stack */
    static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    pthread_mutex_lock(&mu);
    pthread_t th[NTHREAD];
    int started[NTHREAD];
    const uint32_t span = NK / NTHREAD; /* whole regions and whole bytes per thread */
    for (int t = 0; t < NTHREAD; t++) {
        fixjob* j = &jobs[t];
        j->which = which;
        j->k0 = span * (uint32_t)t;
        j->k1 = span * (uint32_t)(t + 1);
        j->bits = bits;
        j->bad = 0;
        for (size_t r = 0; r < sizeof(j->dmax) / sizeof(j->dmax[0]); r++)
            j->dmax[r][0] = j->dmax[r][1] = j->dmax[r][2] = -1.0;
        started[t] = pthread_create(&th[t], NULL, fix_range, j) == 0;
        if (!started[t]) fix_range(j);
    }
    int64_t bad = 0;
    for (int t = 0; t < NTHREAD; t++) {
        if (started[t]) pthread_join(th[t], NULL);
        const fixjob* j = &jobs[t];
        bad += j->bad;
        const uint32_t r0 = j->k0 >> MMC_LIBM_REGSHIFT, r1 = j->k1 >> MMC_LIBM_REGSHIFT;
        for (uint32_t r = r0; r < r1; r++)
            for (int v = 0; v < nval; v++) {
                const double q = j->dmax[r - r0][v];
                float d = (float)q;
                if ((double)d < q) d = nextafterf(d, INFINITY);
                thr[r * (uint32_t)nval + (uint32_t)v] = q < 0.0 ? -1.f : d;
            }
    }
    pthread_mutex_unlock(&mu);
    return bad > 0x7FFFFFFF ? 0x7FFFFFFF : (int32_t)bad;
}
