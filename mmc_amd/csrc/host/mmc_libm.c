/*
 * mmc_libm.c -- agreement tables between the deterministic FP32 functions of
 * include/mmc_fp32math.h and the host C library's expf/logf/sincosf/acosf.
 *
 * The reference CPU tracer draws every scattering length, azimuth and
 * isotropic zenith through the C library (src/mmc_rand_common.h:50-51,
 * src/mmc_mesh.c:1662-1690, src/mmc_raytrace.c:2434-2480), and a one-ulp
 * difference in any of them moves the photon for the rest of its life.  The
 * C library's float functions are not correctly rounded, so no independent
 * implementation reproduces them by construction.  But each of these calls
 * sees only a function of one generator output: u = k / 2^23 with k the 23
 * mantissa bits of xorshift128+ (src/mmc_rand_xorshift128p.c:55-70).  For
 * every k the table below stores how the C library's result differs from the
 * deterministic FP32 evaluation, as a 2-bit code per value:
 *     0 equal, 1 one ulp up (raw bits + 1), 2 one ulp down (raw bits - 1),
 *     3 look the exact values up in the escape list.
 * The device evaluates the FP32 function and applies the code, so its result
 * equals the host C library's on every input the generator can produce.
 *
 * Tables (k = 0 .. 2^23-1):
 *   MMC_FIX_LOG  logf(u + 1e-6f)                        2 bits / k
 *   MMC_FIX_AZ   sincosf((float)(2 pi u))               sin bits 0-1, cos 2-3
 *   MMC_FIX_ISO  t = acosf(2u - 1), sincosf(t)          acos 0-1, sin 2-3, cos 4-5
 *   MMC_FIX_PI   sincosf((float)(pi u))                 sin bits 0-1, cos 2-3
 * The ISO table corrects acos first and then the sin/cos of the corrected
 * angle, so each stage stays within one ulp.
 *
 * Compiled with -mfma (the FP32 functions' explicit fused multiply-adds
 * become single instructions, as on the GPU); callers check that the host
 * CPU has FMA before calling.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "mmc_fp32math.h"
#include "mmc_host.h"

#define NK (1u << 23)
#define NTHREAD 8

static int32_t bitdiff(float ref, float mine) {
    uint32_t a = mmc_fbits(ref), b = mmc_fbits(mine);
    if ((a ^ b) & 0x80000000u) return 99; /* sign differs: escape */
    return (int32_t)(a - b);
}
static uint32_t code_of(int32_t d) { return d == 0 ? 0u : d == 1 ? 1u : d == -1 ? 2u : 3u; }

/* The argument of each table at generator output k, exactly as the kernel
 * forms it. */
static float u_of(uint32_t k) { return mmc_bitsf(0x3F800000u | k) - 1.0f; }

typedef struct {
    int32_t which;
    uint32_t k0, k1;
    uint8_t* out;
    mmc_libm_escape esc[MMC_LIBM_MAXESC];
    int32_t nesc;
} fixjob;

static void add_escape(fixjob* j, uint32_t k, float v0, float v1, float v2) {
    if (j->nesc < MMC_LIBM_MAXESC) {
        mmc_libm_escape* e = &j->esc[j->nesc];
        e->key = (int32_t)(((uint32_t)j->which << 23) | k);
        e->v[0] = mmc_fbits(v0);
        e->v[1] = mmc_fbits(v1);
        e->v[2] = mmc_fbits(v2);
    }
    j->nesc++;
}

static void* fix_range(void* arg) {
    fixjob* j = (fixjob*)arg;
    for (uint32_t k = j->k0; k < j->k1; k++) {
        const float u = u_of(k);
        if (j->which == MMC_FIX_LOG) {
            const float x = u + 1e-6f;
            const float ref = logf(x), mine = mmc_logf(x);
            uint32_t c = code_of(bitdiff(ref, mine));
            if (c == 3) add_escape(j, k, ref, 0.f, 0.f);
            j->out[k >> 2] |= (uint8_t)(c << (2 * (k & 3)));
        } else if (j->which == MMC_FIX_AZ || j->which == MMC_FIX_PI) {
            const float a = (j->which == MMC_FIX_AZ) ? mmc_twopi(u) : mmc_pi(u);
            float s, c, ms, mc;
            sincosf(a, &s, &c);
            mmc_sincosf(a, &ms, &mc);
            uint32_t cs = code_of(bitdiff(s, ms)), cc = code_of(bitdiff(c, mc));
            if (cs == 3 || cc == 3) {
                cs = cc = 3;
                add_escape(j, k, s, c, 0.f);
            }
            j->out[k >> 1] |= (uint8_t)((cs | (cc << 2)) << (4 * (k & 1)));
        } else { /* MMC_FIX_ISO */
            const float x = 2.f * u - 1.f;
            const float t = acosf(x), mt = mmc_acosf(x);
            uint32_t ct = code_of(bitdiff(t, mt));
            float s, c, ms, mc;
            sincosf(t, &s, &c);
            mmc_sincosf(t, &ms, &mc);
            uint32_t cs = code_of(bitdiff(s, ms)), cc = code_of(bitdiff(c, mc));
            if (ct == 3 || cs == 3 || cc == 3) {
                ct = cs = cc = 3;
                add_escape(j, k, t, s, c);
            }
            j->out[k] = (uint8_t)(ct | (cs << 2) | (cc << 4));
        }
    }
    return NULL;
}

int64_t mmc_host_libm_table_bytes(int32_t which) {
    switch (which) {
    case MMC_FIX_LOG: return NK / 4;
    case MMC_FIX_AZ:
    case MMC_FIX_PI: return NK / 2;
    case MMC_FIX_ISO: return NK;
    default: return -1;
    }
}

int32_t mmc_host_libm_table(int32_t which, uint8_t* out, mmc_libm_escape* esc, int32_t maxesc) {
    const int64_t nbytes = mmc_host_libm_table_bytes(which);
    if (nbytes < 0 || !out) return -1;
    memset(out, 0, (size_t)nbytes);
    fixjob jobs[NTHREAD];
    pthread_t th[NTHREAD];
    int started[NTHREAD];
    const uint32_t span = NK / NTHREAD; /* a multiple of 4: threads never share a byte */
    for (int t = 0; t < NTHREAD; t++) {
        jobs[t].which = which;
        jobs[t].k0 = span * (uint32_t)t;
        jobs[t].k1 = (t == NTHREAD - 1) ? NK : span * (uint32_t)(t + 1);
        jobs[t].out = out;
        jobs[t].nesc = 0;
        started[t] = pthread_create(&th[t], NULL, fix_range, &jobs[t]) == 0;
        if (!started[t]) fix_range(&jobs[t]);
    }
    int32_t n = 0;
    for (int t = 0; t < NTHREAD; t++) {
        if (started[t]) pthread_join(th[t], NULL);
        for (int i = 0; i < jobs[t].nesc && i < MMC_LIBM_MAXESC; i++) {
            if (n < maxesc && esc) esc[n] = jobs[t].esc[i];
            n++;
        }
        if (jobs[t].nesc > MMC_LIBM_MAXESC) n += jobs[t].nesc - MMC_LIBM_MAXESC;
    }
    return n;
}
