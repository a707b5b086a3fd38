/*
 * mmc_oracle.c -- CPU restatement of the fangq/mmc photon-transport path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library; the product never does.
 *
 * This is a clean-room restatement written from the reference's semantics.
 * It follows the reference CPU build (`sse` target: -DMMC_USE_SSE -msse4.1,
 * glibc libm, no -ffast-math, commons/Makefile_common.mk:46-49) operation by
 * operation, including the order of floating-point additions, because the
 * parity tests compare single photon trajectories.  Build with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Citations are relative to the reference root.
 */
#define _GNU_SOURCE
#include "mmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Elementary functions.  The reference CPU build calls glibc's expf, logf,
 * sincosf/sinf/cosf and acosf (no -ffast-math), and double-precision log for
 * the Gaussian sources.  Scattering lengths, azimuths and isotropic zeniths
 * are functions of one generator output, and the GPU reproduces glibc's
 * values on all 2^23 of them, so both builds call glibc there.
 * ORC_PORTABLE_MATH -- the GPU's arithmetic, for bit-for-bit comparison of
 * single photons -- differs from the default build (the reference's
 * arithmetic) only where an argument is not a function of one draw: expf is
 * the deterministic FP32 routine of include/mmc_fp32math.h, sin/cos of such
 * angles are correctly rounded through double precision (O_SINCOSF_C), and
 * the RF phase rotation, which exists only in the reference's GPU code, uses
 * the FP32 routine (O_SINCOSF_RF). */
#ifdef ORC_PORTABLE_MATH
#include "mmc_fp32math.h"
static inline void orc_sincos_cr(float x, float* s, float* c) {
    double sd, cd;
    mmc_sincos_d(x, &sd, &cd);
    *s = (float)sd;
    *c = (float)cd;
}
#define O_EXPF(x) mmc_expf(x)
#define O_SINCOSF_C(x, s, c) orc_sincos_cr(x, s, c)
#define O_SINCOSF_RF(x, s, c) mmc_sincosf(x, s, c)
#else
#define O_EXPF(x) expf(x)
#define O_SINCOSF_C(x, s, c) sincosf(x, s, c)
#define O_SINCOSF_RF(x, s, c) sincosf(x, s, c)
#endif

#define EPS        1e-6f              /* src/mmc_rand_common.h:46 */
#define TWO_PI     (M_PI * 2.0)       /* src/mmc_rand_common.h:45 (double) */
#define R_C0       3.335640951981520e-12f  /* src/mmc_const.h:45 */
#define MAX_TRIAL  3                  /* src/mmc_raytrace.h:44 */
#define FIX_PHOTON 1e-3f              /* src/mmc_raytrace.h:45 */
#define R_MIN_MUS  1e9f               /* src/mmc_mesh.h:72 */
#define UNDEF_F    (3.40282347e+38F)  /* MMC_UNDEFINED, src/mmc_mesh.h:65 */
#define VERY_BIG   1e30f              /* src/mmc_mesh.h:74 */

/* local face / edge tables: src/mmc_raytrace.c:75-152, src/mmc_mesh.c:59-103 */
static const int O_fc[4][3] = {{0, 4, 2}, {3, 5, 4}, {2, 5, 1}, {1, 3, 0}};
static const int O_nc[4][3] = {{3, 0, 1}, {3, 1, 2}, {2, 0, 3}, {1, 0, 2}};
static const int O_out[4][3] = {{0, 3, 1}, {3, 2, 1}, {0, 2, 3}, {0, 1, 2}};
static const int O_facemap[4] = {2, 0, 1, 3};
static const int O_ifacemap[4] = {1, 2, 0, 3};
static const int O_faceorder[5] = {1, 3, 2, 0, -1};
static const int O_ifaceorder[4] = {3, 0, 2, 1};
static const int O_pairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
/* facelist used by the face-neighbour builder, src/mmc_highorder.cpp:50 */
static const int O_facelist[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

/* ------------------------------------------------------------------------ */
/* RNG: xorshift128+ (src/mmc_rand_xorshift128p.c:55-83)                     */
/* ------------------------------------------------------------------------ */

float orc_rand_next(uint64_t t[2]) {
    uint64_t s1 = t[0];
    const uint64_t s0 = t[1];
    t[0] = s0;
    s1 ^= s1 << 23;
    t[1] = s1 ^ s0 ^ (s1 >> 18) ^ (s0 >> 5);
    uint64_t r = t[1] + s0;
    uint32_t lo = (uint32_t)r;                 /* u[0]: low word on little-endian */
    uint32_t bits = 0x3F800000U | (lo >> 9);
    float f;
    memcpy(&f, &bits, 4);
    return f - 1.0f;
}

void orc_rand_seed(const uint32_t seed[4], uint64_t t[2]) {
    t[0] = (uint64_t)seed[0] << 32 | seed[1];
    t[1] = (uint64_t)seed[2] << 32 | seed[3];
}

/* src/mmc_rand_common.h:50-51 */
float orc_rand_scatlen(uint64_t t[2]) {
    return -logf(orc_rand_next(t) + EPS);
}

#ifdef ORC_PORTABLE_MATH
/* ------------------------------------------------------------------------ */
/* Check of the GPU's agreement tables (mmc_host_libm_table), restating the  */
/* kernel's rule: the double evaluation of mmc_fp32math.h rounded to float  */
/* f, replaced by its far neighbour o when the double lies within D |o - f|  */
/* of the midpoint (D: the region's threshold) and the table bit is set.     */
/* Applied over every generator output k; returns the number of values that  */
/* differ from glibc's logf / sincosf / acosf.                               */
/* ------------------------------------------------------------------------ */
static float orc_lm(double e, float D, const uint8_t* bits, uint32_t bitpos) {
    float f = (float)e;
    if (e == (double)f) return f;
    const float o = mmc_fnext(f, e);
    const double fd = f, od = o;
    if (fabs(e - 0.5 * (fd + od)) <= (double)D * fabs(od - fd) && ((bits[bitpos >> 3] >> (bitpos & 7)) & 1u))
        return o;
    return f;
}
static int orc_same(float a, float b) { return memcmp(&a, &b, 4) == 0; }

/* number of generator outputs whose draw of table `which` (0 log, 1 azimuth)
 * lies within its region's threshold of a rounding midpoint, i.e. reads the
 * table on the GPU */
int64_t orc_libm_lookups(int which, const float* thr) {
    int64_t n = 0;
    for (uint32_t k = 0; k < (1u << 23); k++) {
        uint32_t ub = 0x3F800000u | k;
        float u;
        memcpy(&u, &ub, 4);
        u -= 1.0f;
        const uint32_t reg = k >> 16;
        double e[2];
        int nv = 1;
        float D[2] = {thr[reg], 0.f};
        if (which == 0) {
            e[0] = mmc_log_d(u + EPS);
        } else {
            mmc_sincos_d((float)(TWO_PI * u), &e[0], &e[1]);
            nv = 2;
            D[0] = thr[2 * reg];
            D[1] = thr[2 * reg + 1];
        }
        int near = 0;
        for (int v = 0; v < nv; v++) {
            const float f = (float)e[v];
            if (e[v] == (double)f) continue;
            const double fd = f, od = mmc_fnext(f, e[v]);
            near |= fabs(e[v] - 0.5 * (fd + od)) <= (double)D[v] * fabs(od - fd);
        }
        n += near;
    }
    return n;
}

/* largest relative difference of the double evaluations from glibc's double
 * log / sin / cos / acos over every generator output: 0 log(u + EPS), 1 sin
 * and cos of (float)(2 pi u), 2 acos(2u - 1), 3 sin and cos of (float)(pi u) */
double orc_libm_dd_maxrel(int which) {
    double mx = 0.0;
    for (uint32_t k = 0; k < (1u << 23); k++) {
        uint32_t ub = 0x3F800000u | k;
        float u;
        memcpy(&u, &ub, 4);
        u -= 1.0f;
        double e[2], r[2];
        int nv = 1;
        if (which == 0) {
            e[0] = mmc_log_d(u + EPS);
            r[0] = log((double)(u + EPS));
        } else if (which == 2) {
            e[0] = mmc_acos_d(2.f * u - 1.f);
            r[0] = acos((double)(2.f * u - 1.f));
        } else {
            const float a = (which == 1) ? (float)(TWO_PI * u) : (float)(M_PI * u);
            mmc_sincos_d(a, &e[0], &e[1]);
            r[0] = sin((double)a);
            r[1] = cos((double)a);
            nv = 2;
        }
        for (int v = 0; v < nv; v++)
            if (r[v] != 0.0) {
                const double d = fabs(e[v] - r[v]) / fabs(r[v]);
                if (d > mx) mx = d;
            }
    }
    return mx;
}

int64_t orc_libm_verify(int which, const uint8_t* bits, const float* thr) {
    int64_t bad = 0;
    for (uint32_t k = 0; k < (1u << 23); k++) {
        uint32_t ub = 0x3F800000u | k;
        float u;
        memcpy(&u, &ub, 4);
        u -= 1.0f;
        const uint32_t reg = k >> 16;
        if (which == 0) {                        /* logf(u + EPS): 1 bit per k */
            const float y = orc_lm(mmc_log_d(u + EPS), thr[reg], bits, k);
            bad += !orc_same(y, logf(u + EPS));
        } else if (which == 1 || which == 3) {   /* sincosf((float)(2 pi u)) or (float)(pi u): 2 bits */
            const float a = (which == 1) ? (float)(TWO_PI * u) : (float)(M_PI * u);
            double es, ec;
            float gs, gc;
            mmc_sincos_d(a, &es, &ec);
            const float s = orc_lm(es, thr[2 * reg], bits, 2 * k);
            const float c = orc_lm(ec, thr[2 * reg + 1], bits, 2 * k + 1);
            sincosf(a, &gs, &gc);
            bad += !orc_same(s, gs) + !orc_same(c, gc);
        } else {                                 /* t = acosf(2u - 1), sincosf(t): 4 bits */
            const float t = orc_lm(mmc_acos_d(2.f * u - 1.f), thr[3 * reg], bits, 4 * k);
            double es, ec;
            float gs, gc;
            mmc_sincos_d(t, &es, &ec);
            const float s = orc_lm(es, thr[3 * reg + 1], bits, 4 * k + 1);
            const float c = orc_lm(ec, thr[3 * reg + 2], bits, 4 * k + 2);
            const float gt = acosf(2.f * u - 1.f);
            sincosf(gt, &gs, &gc);
            bad += !orc_same(t, gt) + !orc_same(s, gs) + !orc_same(c, gc);
        }
    }
    return bad;
}
#endif

/* ------------------------------------------------------------------------ */
/* small vector helpers with the reference's evaluation order               */
/* (src/mmc_mesh.h:194-357)                                                  */
/* ------------------------------------------------------------------------ */

typedef struct { float x, y, z; } v3;

static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 diff3(v3 a, v3 b) { v3 r = {b.x - a.x, b.y - a.y, b.z - a.z}; return r; }
static inline v3 cross3(v3 a, v3 b) {
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
/* vec_mult_add(a,b,sa,sb): sb*b + sa*a (src/mmc_mesh.h:212-216) */
static inline v3 mult_add(v3 a, v3 b, float sa, float sb) {
    v3 r = {sb * b.x + sa * a.x, sb * b.y + sa * a.y, sb * b.z + sa * a.z};
    return r;
}
static inline v3 node_at(const orc_mesh* m, int id1) {
    const float* p = m->node + 3 * (size_t)(id1 - 1);
    v3 r = {p[0], p[1], p[2]};
    return r;
}
/* SSE4.1 _mm_dp_ps(a,b,0x7F): (ax*bx + ay*by) + (az*bz + 0)  (src/mmc_raytrace.c:264) */
static inline float dpps(v3 a, v3 b) {
    float t2 = a.x * b.x + a.y * b.y;
    float t3 = a.z * b.z + 0.0f;
    return t2 + t3;
}
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* ------------------------------------------------------------------------ */
/* mesh preparation                                                          */
/* ------------------------------------------------------------------------ */

/* mesh_getvolume, src/mmc_mesh.c:910-950: signed volume, swap nodes 3/4 when
 * negative, nodal volume = 1/4 of each non-void element. */
void orc_getvolume(int nn, const float* node, int ne, int* elem, const int* type,
                   float* evol, float* nvol) {
    memset(nvol, 0, sizeof(float) * (size_t)nn);
    for (int i = 0; i < ne; i++) {
        int* ee = elem + 4 * (size_t)i;
        const float* n0 = node + 3 * (size_t)(ee[0] - 1);
        const float* n1 = node + 3 * (size_t)(ee[1] - 1);
        const float* n2 = node + 3 * (size_t)(ee[2] - 1);
        const float* n3 = node + 3 * (size_t)(ee[3] - 1);
        float dx = n2[0] - n3[0], dy = n2[1] - n3[1], dz = n2[2] - n3[2];
        float v = n1[0] * (n2[1] * n3[2] - n2[2] * n3[1])
                  - n1[1] * (n2[0] * n3[2] - n2[2] * n3[0])
                  + n1[2] * (n2[0] * n3[1] - n2[1] * n3[0]);
        v += -n0[0] * ((n2[1] * n3[2] - n2[2] * n3[1]) + n1[1] * dz - n1[2] * dy);
        v += +n0[1] * ((n2[0] * n3[2] - n2[2] * n3[0]) + n1[0] * dz - n1[2] * dx);
        v += -n0[2] * ((n2[0] * n3[1] - n2[1] * n3[0]) + n1[0] * dy - n1[1] * dx);
        v = -v;
        if (v < 0.f) {
            int e1 = ee[3];
            ee[3] = ee[2];
            ee[2] = e1;
            v = -v;
        }
        v *= (1.f / 6.f);
        evol[i] = v;
        if (type[i] == 0) continue;
        for (int j = 0; j < 4; j++) nvol[ee[j] - 1] += v * 0.25f;
    }
}

/* mesh_getfacenb, src/mmc_highorder.cpp:128-159: facenb[i*4+j] is the element
 * sharing face facelist[j] of element i (0 when on the exterior).  Restated
 * with a sort over (sorted triple, owner) records instead of a hash map; the
 * result is the same pairing because an interior face is shared by exactly
 * two elements. */
typedef struct { int a, b, c; int slot; } orc_face;
static int cmp_face(const void* x, const void* y) {
    const orc_face* f = (const orc_face*)x;
    const orc_face* g = (const orc_face*)y;
    if (f->a != g->a) return f->a < g->a ? -1 : 1;
    if (f->b != g->b) return f->b < g->b ? -1 : 1;
    if (f->c != g->c) return f->c < g->c ? -1 : 1;
    return f->slot < g->slot ? -1 : (f->slot > g->slot);
}
void orc_getfacenb(int ne, const int* elem, int* facenb) {
    size_t nfc = (size_t)ne * 4;
    orc_face* fl = (orc_face*)malloc(sizeof(orc_face) * nfc);
    for (int i = 0; i < ne; i++)
        for (int j = 0; j < 4; j++) {
            int v[3] = {elem[4 * i + O_facelist[j][0]], elem[4 * i + O_facelist[j][1]],
                        elem[4 * i + O_facelist[j][2]]};
            /* sort 3 */
            if (v[0] > v[1]) { int t = v[0]; v[0] = v[1]; v[1] = t; }
            if (v[1] > v[2]) { int t = v[1]; v[1] = v[2]; v[2] = t; }
            if (v[0] > v[1]) { int t = v[0]; v[0] = v[1]; v[1] = t; }
            orc_face* f = fl + 4 * (size_t)i + j;
            f->a = v[0]; f->b = v[1]; f->c = v[2]; f->slot = 4 * i + j;
        }
    qsort(fl, nfc, sizeof(orc_face), cmp_face);
    memset(facenb, 0, sizeof(int) * nfc);
    for (size_t k = 0; k + 1 < nfc; k++) {
        if (fl[k].a == fl[k + 1].a && fl[k].b == fl[k + 1].b && fl[k].c == fl[k + 1].c) {
            facenb[fl[k].slot] = fl[k + 1].slot / 4 + 1;
            facenb[fl[k + 1].slot] = fl[k].slot / 4 + 1;
            k++;
        }
    }
    free(fl);
}

/* tracer_build, src/mmc_mesh.c:1496-1602.  planes: per element 16 floats
 * {nx[4], ny[4], nz[4], d[4]}, face j spanned by out[j], unit outward normal
 * with Rn2 = 1.f/sqrt(double).  Plucker: d = n1 - n0, m = n0 x n1 per edge. */
void orc_tracer_build(int ne, const float* node, const int* elem,
                      float* planes, float* pd, float* pm) {
    for (int i = 0; i < ne; i++) {
        const int* ee = elem + 4 * (size_t)i;
        if (pd && pm) {
            for (int j = 0; j < 6; j++) {
                const float* a = node + 3 * (size_t)(ee[O_pairs[j][0]] - 1);
                const float* b = node + 3 * (size_t)(ee[O_pairs[j][1]] - 1);
                v3 A = {a[0], a[1], a[2]}, B = {b[0], b[1], b[2]};
                v3 d = diff3(A, B), mm = cross3(A, B);
                float* od = pd + 18 * (size_t)i + 3 * j;
                float* om = pm + 18 * (size_t)i + 3 * j;
                od[0] = d.x; od[1] = d.y; od[2] = d.z;
                om[0] = mm.x; om[1] = mm.y; om[2] = mm.z;
            }
        }
        if (planes) {
            float* vn = planes + 16 * (size_t)i;
            for (int j = 0; j < 4; j++) {
                const float* a = node + 3 * (size_t)(ee[O_out[j][0]] - 1);
                const float* b = node + 3 * (size_t)(ee[O_out[j][1]] - 1);
                const float* c = node + 3 * (size_t)(ee[O_out[j][2]] - 1);
                v3 A = {a[0], a[1], a[2]}, B = {b[0], b[1], b[2]}, C = {c[0], c[1], c[2]};
                v3 AB = diff3(A, B), AC = diff3(A, C);
                v3 N = cross3(AB, AC);
                float Rn2 = 1.f / sqrt(dot3(N, N));
                N.x = Rn2 * N.x; N.y = Rn2 * N.y; N.z = Rn2 * N.z;
                vn[j] = N.x;
                vn[j + 4] = N.y;
                vn[j + 8] = N.z;
                vn[j + 12] = dot3(N, A);
            }
        }
    }
}

/* mesh_barycentric, src/mmc_mesh.c:1170-1206 */
int orc_barycentric(const orc_mesh* m, int e0, const float srcpos[3], float bary[4]) {
    const int* ee = m->elem + 4 * (size_t)(e0 - 1);
    v3 S0 = {srcpos[0], srcpos[1], srcpos[2]};
    float s = 0.f;
    for (int i = 0; i < 4; i++) {
        v3 A = node_at(m, ee[O_out[i][0]]), B = node_at(m, ee[O_out[i][1]]),
           C = node_at(m, ee[O_out[i][2]]);
        v3 AB = diff3(A, B), AC = diff3(A, C), S = diff3(A, S0);
        v3 N = cross3(AB, AC);
        bary[O_facemap[i]] = -dot3(S, N);
    }
    for (int i = 0; i < 4; i++) {
        if (bary[i] < 0.f) return 1;
        s += bary[i];
    }
    for (int i = 0; i < 4; i++) bary[i] /= s;
    return 0;
}

/* mesh_initelem, src/mmc_mesh.c:1060-1090: first element (index order) whose
 * bounding box holds the source and whose barycentric weights are >= 0. */
int orc_initelem(const orc_mesh* m, const float srcpos[3], float bary[4]) {
    for (int i = 0; i < m->ne; i++) {
        double pmin[3] = {VERY_BIG, VERY_BIG, VERY_BIG}, pmax[3] = {-VERY_BIG, -VERY_BIG, -VERY_BIG};
        const int* ee = m->elem + 4 * (size_t)i;
        for (int j = 0; j < 4; j++) {
            const float* p = m->node + 3 * (size_t)(ee[j] - 1);
            for (int k = 0; k < 3; k++) {
                pmin[k] = p[k] < pmin[k] ? p[k] : pmin[k];
                pmax[k] = p[k] > pmax[k] ? p[k] : pmax[k];
            }
        }
        if (srcpos[0] <= pmax[0] && srcpos[0] >= pmin[0] && srcpos[1] <= pmax[1] &&
            srcpos[1] >= pmin[1] && srcpos[2] <= pmax[2] && srcpos[2] >= pmin[2]) {
            if (orc_barycentric(m, i + 1, srcpos, bary) == 0) return i + 1;
        }
    }
    return 0;
}

/* mesh_getreff, src/mmc_mesh.c:2305-2335 */
double orc_getreff(double n_in, double n_out) {
    double oc = asin(1.0 / n_in);
    const double count = 1000.0;
    const double ostep = (M_PI / (2.0 * count));
    double r_phi = 0.0, r_j = 0.0;
    for (int i = 0; i < count; i++) {
        double o = i * ostep, coso = cos(o), r_fres, tmp, cosop;
        if (o < oc) {
            cosop = n_in * sin(o);
            cosop = sqrt(1. - cosop * cosop);
            tmp = (n_in * cosop - n_out * coso) / (n_in * cosop + n_out * coso);
            r_fres = 0.5 * tmp * tmp;
            tmp = (n_in * coso - n_out * cosop) / (n_in * coso + n_out * cosop);
            r_fres += 0.5 * tmp * tmp;
        } else {
            r_fres = 1.f;
        }
        r_phi += 2.0 * sin(o) * coso * r_fres;
        r_j += 3.0 * sin(o) * coso * coso * r_fres;
    }
    r_phi *= ostep;
    r_j *= ostep;
    return (r_phi + r_j) / (2.0 - r_phi + r_j);
}

/* tracer_prep surface-node correction, src/mmc_mesh.c:1344-1386.  Uses the
 * raw face-neighbour table (0 = exterior), i.e. before exterior numbering. */
void orc_surface_nvol(int nn, int ne, const int* elem, const int* type, const int* facenb_raw,
                      const float* med, int prop, int isreflect, float* nvol) {
    float* Reff = (float*)calloc((size_t)prop + 1, sizeof(float));
    if (isreflect) {
        for (int i = 1; i <= prop; i++) {
            for (int j = 1; j < i; j++) {
                if (med[4 * j + 3] == med[4 * i + 3]) { Reff[i] = Reff[j]; break; }
            }
            if (Reff[i] == 0.f) Reff[i] = (float)orc_getreff(med[4 * i + 3], med[3]);
        }
    }
    for (int i = 0; i < ne; i++) {
        const int* ee = elem + 4 * (size_t)i;
        const int* enb = facenb_raw + 4 * (size_t)i;
        for (int j = 0; j < 4; j++) {
            if (enb[j] == 0) {
                for (int k = 0; k < 3; k++) {
                    int nid = ee[O_out[O_ifaceorder[j]][k]] - 1;
                    if (nvol[nid] > 0.f && type[i] >= 0)
                        nvol[nid] *= -(2.f / (1.0 + Reff[type[i]]));
                }
            }
        }
    }
    free(Reff);
    for (int i = 0; i < nn; i++)
        if (nvol[i] < 0.f) nvol[i] = -nvol[i];
}

/* mesh_createdualmesh, src/mmc_mesh.c:349-381: the node bounding box
 * widened by EPS, and cfg->dim voxels of edge step covering it */
void orc_grid(int nn, const float* node, float step, float nmin[3], int dim[3]) {
    float lo[3] = {VERY_BIG, VERY_BIG, VERY_BIG}, hi[3] = {-VERY_BIG, -VERY_BIG, -VERY_BIG};
    for (int i = 0; i < nn; i++)
        for (int k = 0; k < 3; k++) {
            float v = node[3 * (size_t)i + k];
            lo[k] = (v < lo[k]) ? v : lo[k];
            hi[k] = (v > hi[k]) ? v : hi[k];
        }
    for (int k = 0; k < 3; k++) {
        lo[k] -= EPS;
        hi[k] += EPS;
        float ext = hi[k] - lo[k];
        dim[k] = (int)(ext / step) + 1;
        nmin[k] = lo[k];
    }
}

/* tracer_prep exterior numbering, src/mmc_mesh.c:1466-1474 */
int orc_number_exterior(int ne, int* facenb) {
    int nf = 0;
    for (size_t i = 0; i < (size_t)ne * 4; i++)
        if (facenb[i] == 0) facenb[i] = -(++nf);
    return nf;
}

/* ------------------------------------------------------------------------ */
/* photon transport                                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    v3 p0, vec, pout;
    float bary0[4];
    int eid, faceid, isend, nexteid;
    float weight, photontimer, slen0, slen, Lmove;
    double Eabsorb;
    unsigned int posidx, oldidx;
    double oldweight;
    float* ppath;
    int pout_undef;
    /* RF: imaginary weight, its coalescing register, and the magnitude */
    float weight_im, wmag;
    double oldweight_im;
} oray;

typedef struct {
    const orc_cfg* cfg;
    const orc_mesh* m;
    double* weight;
    double* weight_im;
    float rtstep;
    int framelen;
    double raytet;
    float rw, rt;              /* replay: the photon's detected weight and time */
    /* floating-point tie cases of the current photon: ray-tetrahedron tests
     * whose exit face was decided by a tie rule (Badouel: two or more faces
     * at the same smallest distance, resolved by maskmap; Plucker: more than
     * one face passing the sign test, resolved by face order), and fixphoton
     * nudges (no face found) */
    int32_t ties, nudges;
    double* srcweight;         /* patterns sharing photons: launched, then absorbed energy per pattern */
} octx;

#define DELTA_MUA 1e-4f        /* src/mmc_mesh.h:73 */

/* mmc_is_fluence_like, src/mmc_raytrace.c:57-59 */
static inline int fluence_like(int ot) { return ot != ORC_OT_ENERGY && ot != ORC_OT_WP && ot != ORC_OT_WL; }

/* replay weights of src/mmc_raytrace.c:1523-1543: with the J/L/P outputs the
 * deposit sees the detected weight times exp(-DELTA_MUA*L), L or
 * L*mus/slen0, added to the attenuated weight */
static inline float replay_cw(const octx* c, float currweight, const oray* r, float mus) {
    int ot = c->cfg->outputtype;
    float cw;
    if (ot == ORC_OT_JACOBIAN) {
        cw = O_EXPF(-DELTA_MUA * r->Lmove);
    } else if (ot == ORC_OT_WL) {
        cw = r->Lmove;
    } else if (ot == ORC_OT_WP) {
        if (r->slen0 < EPS) cw = 1;
        else cw = r->Lmove * mus / r->slen0;
    } else {
        return currweight;
    }
    cw *= c->rw;
    cw += r->weight;
    return cw;
}

/* time gate: the detected time for the weighted path/scattering outputs */
static inline int replay_gate(const octx* c, float photontimer) {
    int ot = c->cfg->outputtype;
    if (ot == ORC_OT_WL || ot == ORC_OT_WP) return (int)(c->rt * c->rtstep);
    return (int)((photontimer - c->cfg->tstart) * c->rtstep);
}

static inline const float* med_of(const orc_mesh* m, int t) { return m->med + 4 * (size_t)t; }

/* deposit helper: #pragma omp atomic on the double output array */
static inline void add_w(double* w, size_t idx, double v) {
#pragma omp atomic
    w[idx] += v;
}

/* a deposit v into value idx of the field; with patterns sharing the photon
 * (srcnum > 1) slot k of the value receives v times the photon's weight in
 * pattern k, the pattern value at its launch pixel
 * (src/mmc_raytrace.c:410-423, 1586-1612) */
static inline void depo_w(const octx* c, const oray* r, size_t idx, double v) {
    const orc_cfg* cfg = c->cfg;
    if (cfg->srcnum > 1) {
        const float* pw = cfg->srcpattern + (size_t)r->posidx * cfg->srcnum;
        for (int k = 0; k < cfg->srcnum; k++) add_w(c->weight, idx * cfg->srcnum + k, (double)((float)v * pw[k]));
    } else {
        add_w(c->weight, idx, v);
    }
}

/* branchless_badouel_raytet, src/mmc_raytrace.c:1414-1709 (element output,
 * mmMCX, single source, no iMMC). */
static float orc_badouel(oray* r, octx* c) {
    const orc_cfg* cfg = c->cfg;
    const orc_mesh* m = c->m;
    float baryx;
    int eid = r->eid - 1;
    const float* N = m->planes + 16 * (size_t)eid;

    r->pout_undef = 1;
    r->faceid = -1;
    r->isend = 0;

    float T[4], S[4];
    for (int i = 0; i < 4; i++) {
        float t = N[i] * r->p0.x;
        t = t + N[4 + i] * r->p0.y;
        t = t + N[8 + i] * r->p0.z;
        t = N[12 + i] - t;
        float s = N[i] * r->vec.x;
        s = s + N[4 + i] * r->vec.y;
        s = s + N[8 + i] * r->vec.z;
        t = t / s;
        T[i] = (s > 0.f) ? t : 1e10f;
        S[i] = s;
    }
    (void)S;
    /* _mm_min_ps(T, movehl(T)) then min_ss of lanes 0,1 */
    float m02 = (T[0] < T[2]) ? T[0] : T[2];
    float m13 = (T[1] < T[3]) ? T[1] : T[3];
    baryx = (m02 < m13) ? m02 : m13;
    /* maskmap: highest tied lane (src/mmc_raytrace.c:152,1467) */
    int faceidx = 4, ntied = 0;
    for (int i = 0; i < 4; i++)
        if (T[i] == baryx) {
            faceidx = i;
            ntied++;
        }
    if (ntied > 1 && baryx < 1e10f) c->ties++;
    r->faceid = O_faceorder[faceidx];

    if (r->faceid >= 0 && baryx >= 0) {
        int type = m->type[eid];
        const float* prop = med_of(m, type);
        float mua = prop[0], mus = prop[1], n = prop[3];
        float rc = n * R_C0;
        float currweight = r->weight;
        r->nexteid = m->facenb[4 * (size_t)eid + r->faceid];
        float dlen = (mus <= EPS) ? R_MIN_MUS : r->slen / mus;
        float Lp0 = baryx;
        r->isend = (Lp0 > dlen);
        r->Lmove = (r->isend) ? dlen : Lp0;

        r->pout.x = r->vec.x * baryx + r->p0.x;
        r->pout.y = r->vec.y * baryx + r->p0.y;
        r->pout.z = r->vec.z * baryx + r->p0.z;
        r->pout_undef = 0;

        if ((int)((r->photontimer + r->Lmove * rc - cfg->tstart) * c->rtstep) > cfg->maxgate - 1) {
            r->faceid = -2;
            r->pout_undef = 1;
            r->Lmove = (cfg->tend - r->photontimer) / (n * R_C0) - 1e-4f;
        }
        float totalloss = O_EXPF(-mua * r->Lmove);
        r->weight *= totalloss;
        const float att = totalloss;
        totalloss = 1.f - totalloss;
        const int rf = cfg->omega > 0.f && !cfg->replayweight;
        r->slen -= r->Lmove * mus;
        currweight = replay_cw(c, currweight, r, mus);
        const float q0x = r->p0.x, q0y = r->p0.y, q0z = r->p0.z;

        {
            int framelen = c->framelen;
            float ww = currweight - r->weight;
            r->photontimer += r->Lmove * rc;
            int g = replay_gate(c, r->photontimer);
            int tshift = (g < cfg->maxgate - 1 ? (g > 0 ? g : 0) : cfg->maxgate - 1) * framelen;
            if (rf) {
                /* the magnitude of the complex weight decays as the real
                 * weight does without modulation */
                float wb = r->wmag;
                r->wmag *= att;
                if (mua > 0.f) r->Eabsorb += wb - r->wmag;
            } else if (mua > 0.f) {
                r->Eabsorb += ww;
            }
            if (fluence_like(cfg->outputtype))
                ww = (mua < EPS) ? (currweight * r->Lmove) : ((currweight - r->weight) / mua);
            /* RF: complex Beer-Lambert over the step and the complex deposit
             * (w_before - w_after) / (mua + i omega n / c0)
             * (src/mmc_core.cl:868-897) */
            float ww_im = 0.f;
            const float kap = rf ? cfg->omega * n * R_C0 : 0.f;
            const float amag2 = mua * mua + kap * kap;
            if (rf && cfg->method != ORC_RT_BLBADOUELGRID) {
                float attr = totalloss < 1.f ? (1.f - totalloss) : 1.f;
                float phase = cfg->omega * n * R_C0 * r->Lmove;
                float sphi, cphi;
                O_SINCOSF_RF(phase, &sphi, &cphi);
                float w0r = currweight, w0i = r->weight_im;
                float nre = attr * (w0r * cphi + w0i * sphi);
                float nim = attr * (-w0r * sphi + w0i * cphi);
                float dwr = w0r - nre, dwi = w0i - nim;
                ww = (amag2 > 0.f) ? (dwr * mua + dwi * kap) / amag2 : (w0r * r->Lmove);
                ww_im = (amag2 > 0.f) ? (dwi * mua - dwr * kap) / amag2 : (w0i * r->Lmove);
                r->weight = nre;
                r->weight_im = nim;
            }

            r->p0.x = r->p0.x + r->vec.x * r->Lmove;
            r->p0.y = r->p0.y + r->vec.y * r->Lmove;
            r->p0.z = r->p0.z + r->vec.z * r->Lmove;

            if (cfg->method == ORC_RT_BLBADOUELGRID) {
                /* dual-grid deposit, src/mmc_raytrace.c:1615-1673: the step
                 * is cut into an even number of segments no longer than half
                 * a voxel; each deposits its share of the step's loss at the
                 * voxel holding its midpoint (voxel index clamped to the grid) */
                int seg = (int)(r->Lmove / cfg->gridstep) + 1;
                seg = (seg << 1);
                float dstep = r->Lmove / seg;
                float segloss = O_EXPF(-mua * dstep);
                float tx = r->vec.x * dstep, ty = r->vec.y * dstep, tz = r->vec.z * dstep;
                float sx = (q0x - cfg->gridmin[0]) + tx * 0.5f;
                float sy = (q0y - cfg->gridmin[1]) + ty * 0.5f;
                float sz = (q0z - cfg->gridmin[2]) + tz * 0.5f;
                float inv = 1.f / cfg->gridstep;
                float frac = (totalloss == 0.f) ? 0.f : (1.f - segloss) / totalloss;
                float w0 = ww;
                /* RF: the complex weight turns by the same phase in every
                 * segment (src/mmc_core.cl:1043-1081) */
                float sr = currweight, si = r->weight_im, scos = 1.f, ssin = 0.f;
                if (rf) O_SINCOSF_RF(cfg->omega * n * R_C0 * dstep, &ssin, &scos);
                const int dx = cfg->griddim[0], dxy = cfg->griddim[0] * cfg->griddim[1];
                for (int i = 0; i < seg; i++) {
                    int ix = (int)(sx * inv), iy = (int)(sy * inv), iz = (int)(sz * inv);
                    ix = ix < 0 ? 0 : (ix >= cfg->griddim[0] ? cfg->griddim[0] - 1 : ix);
                    iy = iy < 0 ? 0 : (iy >= cfg->griddim[1] ? cfg->griddim[1] - 1 : iy);
                    iz = iz < 0 ? 0 : (iz >= cfg->griddim[2] ? cfg->griddim[2] - 1 : iz);
                    unsigned int newidx = (unsigned int)(iz * dxy + iy * dx + ix + tshift);
                    r->oldidx = (r->oldidx == 0xFFFFFFFFu) ? newidx : r->oldidx;
                    float d = w0 * frac, dim = 0.f;
                    if (rf) {
                        float w0r = sr, w0i = si;
                        float nre = segloss * (w0r * scos + w0i * ssin);
                        float nim = segloss * (-w0r * ssin + w0i * scos);
                        sr = nre;
                        si = nim;
                        float dwr = w0r - nre, dwi = w0i - nim;
                        d = (amag2 < EPS) ? (w0r * w0) : (dwr * mua + dwi * kap) / amag2;
                        dim = (amag2 < EPS) ? (w0i * w0) : (dwi * mua - dwr * kap) / amag2;
                    }
                    if (newidx != r->oldidx) {
                        depo_w(c, r, r->oldidx, r->oldweight);
                        if (rf) add_w(c->weight_im, r->oldidx, r->oldweight_im);
                        r->oldidx = newidx;
                        r->oldweight = d;
                        r->oldweight_im = dim;
                    } else {
                        r->oldweight += d;
                        r->oldweight_im += dim;
                    }
                    if (r->faceid == -2 || !r->isend) {
                        depo_w(c, r, newidx, r->oldweight);
                        if (rf) add_w(c->weight_im, newidx, r->oldweight_im);
                        r->oldweight = 0.f;
                        r->oldweight_im = 0.f;
                    }
                    w0 *= segloss;
                    sx = sx + tx;
                    sy = sy + ty;
                    sz = sz + tz;
                }
                if (rf) {
                    r->weight = sr;
                    r->weight_im = si;
                }
            } else if (!cfg->basisorder) {
                unsigned int newidx = (unsigned int)(eid + tshift);
                r->oldidx = (r->oldidx == 0xFFFFFFFFu) ? newidx : r->oldidx;
                if (newidx != r->oldidx) {
                    /* RF: the GPU reference flushes only a register whose real
                     * part is positive (src/mmc_core.cl:905) */
                    if (!rf) {
                        depo_w(c, r, r->oldidx, r->oldweight);
                    } else if (r->oldweight > 0.0) {
                        add_w(c->weight, r->oldidx, r->oldweight);
                        add_w(c->weight_im, r->oldidx, r->oldweight_im);
                    }
                    r->oldidx = newidx;
                    r->oldweight = ww;
                    r->oldweight_im = ww_im;
                } else {
                    r->oldweight += ww;
                    r->oldweight_im += ww_im;
                }
                if (r->faceid == -2 || !r->isend) {
                    if (rf) add_w(c->weight, newidx, r->oldweight);
                    else depo_w(c, r, newidx, r->oldweight);
                    if (rf) add_w(c->weight_im, newidx, r->oldweight_im);
                    r->oldweight = 0.f;
                    r->oldweight_im = 0.f;
                }
            } else {
                /* basisorder=1 under BLBadouel: w/3 to the exit-face nodes
                 * (src/mmc_raytrace.c:1674-1690) */
                const int* ee = m->elem + 4 * (size_t)eid;
                ww *= 1.f / 3.f;
                for (int i = 0; i < 3; i++) depo_w(c, r, (size_t)(ee[O_out[faceidx][i]] - 1 + tshift), ww);
            }
        }
    }
    c->raytet++;
    if (r->faceid == -2) return 0.f;
    return r->slen;
}

/* plucker_raytet, src/mmc_raytrace.c:227-508 (SSE4.1 inner products) */
static float orc_plucker(oray* r, octx* c) {
    const orc_cfg* cfg = c->cfg;
    const orc_mesh* m = c->m;
    int eid = r->eid - 1, faceidx = -1;
    float w[6];
    float baryout[4] = {0.f, 0.f, 0.f, 0.f};
    float Lp0 = 0.f, dlen = 0.f;

    r->faceid = -1;
    r->isend = 0;
    r->Lmove = 0.f;
    v3 p1 = {r->p0.x + r->vec.x, r->p0.y + r->vec.y, r->p0.z + r->vec.z};
    v3 pcrx = cross3(r->p0, p1);
    const int* ee = m->elem + 4 * (size_t)eid;
    int type = m->type[eid];
    const float* prop = med_of(m, type);
    float mua = prop[0], mus = prop[1], n = prop[3];
    float rc = n * R_C0;
    float currweight = r->weight;

    for (int i = 0; i < 6; i++) {
        const float* d = m->pd + 18 * (size_t)eid + 3 * i;
        const float* mm = m->pm + 18 * (size_t)eid + 3 * i;
        v3 D = {d[0], d[1], d[2]}, M = {mm[0], mm[1], mm[2]};
        w[i] = dpps(r->vec, M) + dpps(pcrx, D);
    }
    r->pout_undef = 1;
    {
        /* faces passing the sign test (a tie when more than one does) */
        int npass = 0;
        for (int i = 0; i < 4; i++) {
            float wa = w[O_fc[i][0]], wb = (i >= 2) ? -w[O_fc[i][1]] : w[O_fc[i][1]], wc = w[O_fc[i][2]];
            uint32_t a = fbits(wa), b = fbits(wb), cc = fbits(wc);
            npass += ((a & 0x80000000u) & (b & 0x80000000u) & (cc ^ 0x80000000u)) != 0;
        }
        if (npass > 1) c->ties++;
    }
    for (int i = 0; i < 4; i++) {
        if (i >= 2) w[O_fc[i][1]] = -w[O_fc[i][1]];
        uint32_t a = fbits(w[O_fc[i][0]]), b = fbits(w[O_fc[i][1]]), cc = fbits(w[O_fc[i][2]]);
        if ((a & 0x80000000u) & (b & 0x80000000u) & (cc ^ 0x80000000u)) {
            float Rv = 1.f / (-w[O_fc[i][0]] - w[O_fc[i][1]] + w[O_fc[i][2]]);
            baryout[O_nc[i][0]] = -w[O_fc[i][0]] * Rv;
            baryout[O_nc[i][1]] = -w[O_fc[i][1]] * Rv;
            baryout[O_nc[i][2]] = w[O_fc[i][2]] * Rv;
            v3 P1 = node_at(m, ee[O_nc[i][0]]), P2 = node_at(m, ee[O_nc[i][1]]),
               P3 = node_at(m, ee[O_nc[i][2]]);
            float b1 = baryout[O_nc[i][0]], b2 = baryout[O_nc[i][1]], b3 = baryout[O_nc[i][2]];
            r->pout.x = b1 * P1.x + b2 * P2.x + b3 * P3.x;
            r->pout.y = b1 * P1.y + b2 * P2.y + b3 * P3.y;
            r->pout.z = b1 * P1.z + b2 * P2.z + b3 * P3.z;
            r->pout_undef = 0;
            {
                float dx = r->pout.x - r->p0.x, dy = r->pout.y - r->p0.y, dz = r->pout.z - r->p0.z;
                Lp0 = sqrtf(dx * dx + dy * dy + dz * dz);
            }
            dlen = (mus <= EPS) ? R_MIN_MUS : r->slen / mus;
            faceidx = i;
            r->faceid = O_faceorder[i];
            r->nexteid = m->facenb[4 * (size_t)eid + r->faceid];
            r->isend = (Lp0 > dlen);
            r->Lmove = (r->isend) ? dlen : Lp0;
            break;
        }
    }
    c->raytet++;

    if (!r->pout_undef) {
        if ((int)((r->photontimer + r->Lmove * rc - cfg->tstart) * c->rtstep) >=
            (int)((cfg->tend - cfg->tstart) * c->rtstep)) {
            r->faceid = -2;
            r->pout_undef = 1;
            r->Lmove = (cfg->tend - r->photontimer) / (n * R_C0) - 1e-4f;
        }
        r->weight *= O_EXPF(-mua * r->Lmove);
        r->slen -= r->Lmove * mus;
        currweight = replay_cw(c, currweight, r, mus);
        r->p0.x += r->Lmove * r->vec.x;
        r->p0.y += r->Lmove * r->vec.y;
        r->p0.z += r->Lmove * r->vec.z;

        if (!cfg->basisorder) {
            float ww = currweight - r->weight;
            r->Eabsorb += ww;
            if (fluence_like(cfg->outputtype))
                ww = (mua < EPS) ? (currweight * r->Lmove) : ((currweight - r->weight) / mua);
            r->photontimer += r->Lmove * rc;
            int g = replay_gate(c, r->photontimer);
            int tshift = (g < cfg->maxgate - 1 ? (g > 0 ? g : 0) : cfg->maxgate - 1) * m->ne;
            depo_w(c, r, (size_t)eid + tshift, ww);
        } else if (Lp0 > EPS) {
            r->photontimer += r->Lmove * rc;
            float ww = currweight - r->weight;
            if (mua > 0.f || fluence_like(cfg->outputtype)) {
                float ratio = r->Lmove / Lp0;
                if (mua > 0.f) r->Eabsorb += ww;
                if (fluence_like(cfg->outputtype))
                    ww = (mua < EPS) ? (currweight * r->Lmove) : ((currweight - r->weight) / mua);
                int g = replay_gate(c, r->photontimer);
                int tshift = (g < cfg->maxgate - 1 ? (g > 0 ? g : 0) : cfg->maxgate - 1) * m->nn;
                ww *= 0.5f;
                if (r->isend)
                    for (int i = 0; i < 4; i++)
                        baryout[i] = (1.f - ratio) * r->bary0[i] + ratio * baryout[i];
                for (int i = 0; i < 4; i++)
                    depo_w(c, r, (size_t)(ee[i] - 1 + tshift), ww * (r->bary0[i] + baryout[i]));
            }
            if (r->isend) {
                memcpy(r->bary0, baryout, sizeof(float) * 4);
            } else if (r->nexteid > 0 && faceidx >= 0) {
                const int* nextenb = m->elem + 4 * (size_t)(r->nexteid - 1);
                memset(r->bary0, 0, sizeof(float) * 4);
                for (int j = 0; j < 3; j++)
                    for (int k = 0; k < 4; k++)
                        if (ee[O_nc[faceidx][j]] == nextenb[k]) {
                            r->bary0[k] = baryout[O_nc[faceidx][j]];
                            break;
                        }
            }
        }
        if (r->faceid == -2) return 0.f;
    }
    return r->slen;
}

static inline float orc_trace(oray* r, octx* c) {
    return (c->cfg->method == ORC_RT_PLUCKER) ? orc_plucker(r, c) : orc_badouel(r, c);
}

/* One tracer call from (p, v) in element eid with remaining scattering
 * length slen and a far time limit; out = {faceid, isend, Lmove, pout x,y,z,
 * defined, nexteid}.  For unit tests of the ray-tetrahedron tests. */
void orc_probe_raytet(const orc_mesh* m, int method, int eid, const float p[3], const float v[3],
                      float slen, float out[8]) {
    orc_cfg cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.method = method;
    cfg.basisorder = (method == ORC_RT_PLUCKER);
    cfg.tstart = 0.f;
    cfg.tend = 1.f;
    cfg.tstep = 1.f;
    cfg.maxgate = 1;
    size_t n = (size_t)(cfg.basisorder ? m->nn : m->ne);
    double* scratch = (double*)calloc(n, sizeof(double));
    octx c = {&cfg, m, scratch, NULL, 1.f, (int)n, 0.0};
    oray r;
    memset(&r, 0, sizeof(r));
    r.p0.x = p[0]; r.p0.y = p[1]; r.p0.z = p[2];
    r.vec.x = v[0]; r.vec.y = v[1]; r.vec.z = v[2];
    r.eid = eid;
    r.weight = 1.f;
    r.slen = slen;
    r.oldidx = 0xFFFFFFFFu;
    r.bary0[0] = r.bary0[1] = r.bary0[2] = r.bary0[3] = 0.25f;
    float px = r.p0.x, py = r.p0.y, pz = r.p0.z;
    orc_trace(&r, &c);
    out[0] = (float)r.faceid;
    out[1] = (float)r.isend;
    out[2] = r.Lmove;
    out[3] = r.pout_undef ? px : r.pout.x;
    out[4] = r.pout_undef ? py : r.pout.y;
    out[5] = r.pout_undef ? pz : r.pout.z;
    out[6] = (float)!r.pout_undef;
    out[7] = (float)r.nexteid;
    free(scratch);
}

/* fixphoton, src/mmc_raytrace.c:200-213 */
static void orc_fixphoton(oray* r, const orc_mesh* m) {
    const int* ee = m->elem + 4 * (size_t)(r->eid - 1);
    v3 c0 = {0.f, 0.f, 0.f};
    for (int i = 0; i < 4; i++) {
        v3 p = node_at(m, ee[i]);
        c0.x = c0.x + p.x; c0.y = c0.y + p.y; c0.z = c0.z + p.z;
    }
    r->p0.x += (c0.x * 0.25f - r->p0.x) * FIX_PHOTON;
    r->p0.y += (c0.y * 0.25f - r->p0.y) * FIX_PHOTON;
    r->p0.z += (c0.z * 0.25f - r->p0.z) * FIX_PHOTON;
}

/* mc_next_scatter, src/mmc_mesh.c:1644-1715 (non-SSE-math build); *mom
 * receives the momentum transfer 1 - cos(theta) (src/mmc_mesh.c:1707-1709) */
static float orc_scatter_mom(float g, float dir[3], uint64_t t[2], float* mom);
float orc_scatter(float g, float dir[3], uint64_t t[2]) {
    float mom;
    return orc_scatter_mom(g, dir, t, &mom);
}

static float orc_scatter_mom(float g, float dir[3], uint64_t t[2], float* mom) {
    float sphi, cphi, tmp0, stheta, ctheta, tmp1;
    float nextslen = orc_rand_scatlen(t);
    tmp0 = (float)(TWO_PI * orc_rand_next(t));
    sincosf(tmp0, &sphi, &cphi);
    if (g > EPS) {
        tmp0 = (1.f - g * g) / (1.f - g + 2.f * g * orc_rand_next(t));
        tmp0 *= tmp0;
        tmp0 = (1.f + g * g - tmp0) / (2.f * g);
        if (tmp0 > 1.f) tmp0 = 1.f;
        if (tmp0 < -1.f) tmp0 = -1.f;
        stheta = (float)sqrt(1.f - tmp0 * tmp0);
        ctheta = tmp0;
    } else {
        float theta = acosf(2.f * orc_rand_next(t) - 1.f);
        sincosf(theta, &stheta, &ctheta);
    }
    float px, py, pz;
    if (dir[2] > -1.f + EPS && dir[2] < 1.f - EPS) {
        tmp0 = 1.f - dir[2] * dir[2];
        tmp1 = 1.f / sqrtf(tmp0);
        tmp1 = stheta * tmp1;
        px = tmp1 * (dir[0] * dir[2] * cphi - dir[1] * sphi) + dir[0] * ctheta;
        py = tmp1 * (dir[1] * dir[2] * cphi + dir[0] * sphi) + dir[1] * ctheta;
        pz = -tmp1 * tmp0 * cphi + dir[2] * ctheta;
    } else {
        px = stheta * cphi;
        py = stheta * sphi;
        pz = (dir[2] > 0.f) ? ctheta : -ctheta;
    }
    *mom = 1.f - ctheta;
    dir[0] = px;
    dir[1] = py;
    dir[2] = pz;
    return nextslen;
}

/* reflectray, src/mmc_raytrace.c:2257-2332 (no iMMC) */
float orc_reflect(const orc_mesh* m, int method, float nout, int isreflect, int isspecular,
                  float dir[3], int oldeid, int* eid, int faceid, uint64_t t[2]) {
    (void)method;
    int offs = (oldeid - 1) << 2;
    int f = O_ifaceorder[faceid];
    const float* N = m->planes + 4 * (size_t)offs;
    v3 pn = {N[f], N[f + 4], N[f + 8]};
    v3 c0 = {dir[0], dir[1], dir[2]};
    float Icos = fabs(dot3(c0, pn));
    float n1 = (oldeid != *eid) ? med_of(m, m->type[oldeid - 1])[3] : nout;
    float n2 = (*eid > 0) ? med_of(m, m->type[*eid - 1])[3] : nout;
    float tmp0 = n1 * n1, tmp1 = n2 * n2;
    float tmp2 = 1.f - tmp0 / tmp1 * (1.f - Icos * Icos);
    float ret = 1.f;
    if (tmp2 > 0.f && !(*eid <= 0 && isreflect == ORC_BC_MIRROR)) {
        float Re = tmp0 * Icos * Icos + tmp1 * tmp2;
        tmp2 = sqrtf(tmp2);
        float Im = 2.f * n1 * n2 * Icos * tmp2;
        float Rtotal = (Re - Im) / (Re + Im);
        Re = tmp1 * Icos * Icos + tmp0 * tmp2 * tmp2;
        Rtotal = (Rtotal + (Re - Im) / (Re + Im)) * 0.5f;
        if (oldeid == *eid) return Rtotal;
        if (orc_rand_next(t) <= Rtotal) {
            c0 = mult_add(pn, c0, -2.f * Icos, 1.f);
            *eid = oldeid;
        } else if (isspecular == 2 && *eid == 0) {
        } else {
            c0 = mult_add(pn, c0, -Icos, 1.f);
            c0 = mult_add(pn, c0, tmp2, n1 / n2);
        }
    } else {
        c0 = mult_add(pn, c0, -2.f * Icos, 1.f);
        *eid = oldeid;
    }
    tmp0 = 1.f / sqrtf(dot3(c0, c0));
    dir[0] = tmp0 * c0.x;
    dir[1] = tmp0 * c0.y;
    dir[2] = tmp0 * c0.z;
    return ret;
}

/* rotatevector, src/mmc_mesh.h:336-357 (non-SSE build) */
static void orc_rotate(v3* d, float stheta, float ctheta, float sphi, float cphi) {
    v3 p;
    float tmp0;
    if (d->z > -1.f + EPS && d->z < 1.f - EPS) {
        tmp0 = 1.f - d->z * d->z;
        float tmp1 = stheta * (1.f / sqrtf(tmp0));
        p.x = tmp1 * (d->x * d->z * cphi - d->y * sphi) + d->x * ctheta;
        p.y = tmp1 * (d->y * d->z * cphi + d->x * sphi) + d->y * ctheta;
        p.z = -tmp1 * tmp0 * cphi + d->z * ctheta;
    } else {
        p.x = stheta * cphi;
        p.y = stheta * sphi;
        p.z = (d->z > 0.f) ? ctheta : -ctheta;
    }
    tmp0 = 1.f / sqrtf(p.x * p.x + p.y * p.y + p.z * p.z);
    d->x = p.x * tmp0;
    d->y = p.y * tmp0;
    d->z = p.z * tmp0;
}

#define ORC_JUST_BELOW_ONE 0.9998f    /* src/mmc_raytrace.c:54 */
#define ORC_CONE_TRIES (1 << 16)
static inline int imin(int a, int b) { return a < b ? a : b; }

/* launchphoton, src/mmc_raytrace.c:2346-2671, every source type.  Where the
 * reference calls the double-precision log it uses logf here, the second
 * draw of the z-Gaussian source comes from the photon's own stream (the
 * reference takes it from a second generator), the cone's rejection loop is
 * bounded, and the line source's direction is normalized.  Returns nonzero
 * if no candidate element encloses the start. */
static int orc_launch(oray* r, octx* c, uint64_t ran[2]) {
    const orc_cfg* cfg = c->cfg;
    const orc_mesh* m = c->m;
    const int st = cfg->srctype;
    const float *P1 = cfg->srcparam1, *P2 = cfg->srcparam2, *S = cfg->srcpos, *D = cfg->srcdir;
    const float focus = D[3];
    v3 origin = {r->p0.x, r->p0.y, r->p0.z};
    int canfocus = 1;
    r->slen = orc_rand_scatlen(ran);
    if (st == ORC_ST_PENCIL) {
        if (r->eid > 0) return 0;
    } else if (st == ORC_ST_PLANAR || st == ORC_ST_PATTERN || st == ORC_ST_PATTERN3D || st == ORC_ST_FOURIER ||
               st == ORC_ST_PENCILARRAY) {
        float rx = orc_rand_next(ran);
        float ry = orc_rand_next(ran);
        if (st == ORC_ST_PATTERN3D) {
            float rz = orc_rand_next(ran);
            int nx = (int)P1[0], ny = (int)P1[1], nz = (int)P1[2];
            r->p0.x = S[0] + rx * P1[0];
            r->p0.y = S[1] + ry * P1[1];
            r->p0.z = S[2] + rz * P1[2];
            r->weight = (cfg->srcnum > 1) ? 1.f
                : cfg->srcpattern[imin((int)(rz * ORC_JUST_BELOW_ONE * nz), nz - 1) * ny * nx
                                  + imin((int)(ry * ORC_JUST_BELOW_ONE * ny), ny - 1) * nx
                                  + imin((int)(rx * ORC_JUST_BELOW_ONE * nx), nx - 1)];
            origin.x += P1[0] * 0.5f;
            origin.y += P1[1] * 0.5f;
            origin.z += P1[2] * 0.5f;
        } else if (st == ORC_ST_PENCILARRAY) {
            r->p0.x = S[0] + floorf(rx * P1[3]) * P1[0] / (P1[3] - 1.f) + floorf(ry * P2[3]) * P2[0] / (P2[3] - 1.f);
            r->p0.y = S[1] + floorf(rx * P1[3]) * P1[1] / (P1[3] - 1.f) + floorf(ry * P2[3]) * P2[1] / (P2[3] - 1.f);
            r->p0.z = S[2] + floorf(rx * P1[3]) * P1[2] / (P1[3] - 1.f) + floorf(ry * P2[3]) * P2[2] / (P2[3] - 1.f);
            r->weight = 1.f;
            origin.x += (P1[0] + P2[0]) * 0.5f;
            origin.y += (P1[1] + P2[1]) * 0.5f;
            origin.z += (P1[2] + P2[2]) * 0.5f;
        } else {
            r->p0.x = S[0] + rx * P1[0] + ry * P2[0];
            r->p0.y = S[1] + rx * P1[1] + ry * P2[1];
            r->p0.z = S[2] + rx * P1[2] + ry * P2[2];
            r->weight = 1.f;
            if (st == ORC_ST_PATTERN) {
                int xsize = (int)P1[3];
                int ysize = (int)P2[3];
                int iy = (int)(ry * ysize), ix = (int)(rx * xsize);
                iy = iy < ysize - 1 ? iy : ysize - 1;
                ix = ix < xsize - 1 ? ix : xsize - 1;
                r->posidx = (unsigned int)(iy * xsize + ix);
                r->weight = (cfg->srcnum == 1) ? cfg->srcpattern[r->posidx] : 1.f;
                /* seed replay of the weighted path / scattering outputs:
                 * the photon's detected weight takes its launch pixel's
                 * pattern value (src/mmc_raytrace.c:2394-2399) */
                if (cfg->replayweight && (cfg->outputtype == ORC_OT_WL || cfg->outputtype == ORC_OT_WP))
                    c->rw *= cfg->srcpattern[r->posidx];
            } else if (st == ORC_ST_FOURIER) {
                float sn, cs;
                O_SINCOSF_C((float)((floorf(P1[3]) * rx + floorf(P2[3]) * ry + P1[3] - floorf(P1[3])) * TWO_PI),
                          &sn, &cs);
                r->weight = (cs * (1.f - P2[3] + floorf(P2[3])) + 1.f) * 0.5f;
            }
            origin.x += (P1[0] + P2[0]) * 0.5f;
            origin.y += (P1[1] + P2[1]) * 0.5f;
            origin.z += (P1[2] + P2[2]) * 0.5f;
        }
    } else if (st == ORC_ST_FOURIERX || st == ORC_ST_FOURIERX2D) {
        float rx = orc_rand_next(ran);
        float ry = orc_rand_next(ran);
        float v2w = P1[3] * (1.f / sqrtf(P1[0] * P1[0] + P1[1] * P1[1] + P1[2] * P1[2]));
        v3 v2 = {v2w * (D[1] * P1[2] - D[2] * P1[1]), v2w * (D[2] * P1[0] - D[0] * P1[2]),
                 v2w * (D[0] * P1[1] - D[1] * P1[0])};
        r->p0.x = S[0] + rx * P1[0] + ry * v2.x;
        r->p0.y = S[1] + rx * P1[1] + ry * v2.y;
        r->p0.z = S[2] + rx * P1[2] + ry * v2.z;
        if (st == ORC_ST_FOURIERX2D) {
            float sa, ca, sb, cb;
            O_SINCOSF_C((float)((P2[0] * rx + P2[2]) * TWO_PI), &sa, &ca);
            O_SINCOSF_C((float)((P2[1] * ry + P2[3]) * TWO_PI), &sb, &cb);
            r->weight = (sa * sb + 1.f) * 0.5f;
        } else {
            float sn, cs;
            O_SINCOSF_C((float)((P2[0] * rx + P2[1] * ry + P2[2]) * TWO_PI), &sn, &cs);
            r->weight = (cs * (1.f - P2[3]) + 1.f) * 0.5f;
        }
        origin.x += (P1[0] + v2.x) * 0.5f;
        origin.y += (P1[1] + v2.y) * 0.5f;
        origin.z += (P1[2] + v2.z) * 0.5f;
    } else if (st == ORC_ST_DISK || st == ORC_ST_RING || st == ORC_ST_GAUSSIAN) {
        float sphi, cphi, phi, r0;
        if (st == ORC_ST_RING && (P1[2] > 0.f || P1[3] > 0.f)) {
            phi = fabsf(P1[2] - P1[3]) * orc_rand_next(ran) + fminf(P1[2], P1[3]);
            O_SINCOSF_C(phi, &sphi, &cphi);
        } else {
            phi = (float)(TWO_PI * orc_rand_next(ran));
            sincosf(phi, &sphi, &cphi);
        }
        /* double-precision log, converted to float for sqrtf
         * (src/mmc_raytrace.c:2444-2449) */
        if (st == ORC_ST_DISK || st == ORC_ST_RING) {
            r0 = sqrtf(orc_rand_next(ran) * fabsf(P1[0] * P1[0] - P1[1] * P1[1]) + P1[1] * P1[1]);
        } else if (fabsf(focus) < 1e-5f || fabsf(P1[1]) < 1e-5f) {
            r0 = sqrtf(-log(orc_rand_next(ran))) * P1[0];
        } else {
            float z0 = (float)(P1[0] * P1[0] * M_PI / P1[1]);     /* Rayleigh range */
            r0 = sqrtf(-log(orc_rand_next(ran)) * (1.f + (focus * focus / (z0 * z0)))) * P1[0];
        }
        if (D[2] > -1.f + EPS && D[2] < 1.f - EPS) {
            float tmp0 = 1.f - D[2] * D[2];
            float tmp1 = r0 / sqrtf(tmp0);
            r->p0.x = S[0] + tmp1 * (D[0] * D[2] * cphi - D[1] * sphi);
            r->p0.y = S[1] + tmp1 * (D[1] * D[2] * cphi + D[0] * sphi);
            r->p0.z = S[2] - tmp1 * tmp0 * cphi;
        } else {
            r->p0.x += r0 * cphi;
            r->p0.y += r0 * sphi;
        }
    } else if (st == ORC_ST_CONE || st == ORC_ST_ISOTROPIC || st == ORC_ST_ARCSINE) {
        float ang, stheta, ctheta, sphi, cphi;
        ang = (float)(TWO_PI * orc_rand_next(ran));
        sincosf(ang, &sphi, &cphi);
        int clamped = 0;
        if (st == ORC_ST_CONE) {
            int k = 0;
            do {
                ang = (P1[1] > 0) ? (float)(TWO_PI * orc_rand_next(ran)) : acosf(2.f * orc_rand_next(ran) - 1.f);
            } while (ang > P1[0] && ++k < ORC_CONE_TRIES);
            if (ang > P1[0]) {
                ang = P1[0];
                clamped = 1;
            }
        } else if (st == ORC_ST_ISOTROPIC) {
            ang = acosf(2.f * orc_rand_next(ran) - 1.f);
        } else {
            ang = (float)(M_PI * orc_rand_next(ran));
        }
        if (clamped)
            O_SINCOSF_C(ang, &stheta, &ctheta);
        else
            sincosf(ang, &stheta, &ctheta);
        orc_rotate(&r->vec, stheta, ctheta, sphi, cphi);
        canfocus = 0;
        if (r->eid > 0) return 0;
    } else if (st == ORC_ST_ZGAUSSIAN) {
        float ang, stheta, ctheta, sphi, cphi;
        ang = (float)(TWO_PI * orc_rand_next(ran));
        sincosf(ang, &sphi, &cphi);
        /* the reference's second draw comes from ran0, a thread-private state
         * that mmc_run_mp never seeds (src/mmc_host.c:227,265,280-283,
         * src/mmc_rand_xorshift128p.c:82-84); the photon's own stream here */
        float u1 = orc_rand_next(ran);
        float u2 = orc_rand_next(ran);
        ang = sqrtf(-2.f * log(u1)) * (1.f - 2.f * u2) * P1[0];
        O_SINCOSF_C(ang, &stheta, &ctheta);
        orc_rotate(&r->vec, stheta, ctheta, sphi, cphi);
        canfocus = 0;
    } else if (st == ORC_ST_HYPERBOLOID) {
        float sphi, cphi;
        float r0 = (float)(TWO_PI * orc_rand_next(ran));
        sincosf(r0, &sphi, &cphi);
        r0 = sqrtf(0.5f * orc_rand_scatlen(ran)) * P1[0];
        float tilt = (P1[2] > 0.f) ? (-P1[1] / P1[2]) : 0.f;
        float bx = -r0 * sphi, by = r0 * cphi, bz = P1[2];
        float nb = 1.f / sqrtf(bx * bx + by * by + bz * bz);
        bx *= nb;
        by *= nb;
        bz *= nb;
        float dpx = r0 * (cphi - tilt * sphi);
        float dpy = r0 * (sphi + tilt * cphi);
        if (D[2] > -1.f + EPS && D[2] < 1.f - EPS) {
            float sin_t = sqrtf(1.f - D[2] * D[2]);
            float isin = 1.f / sin_t;
            float cpb = D[0] * isin, spb = D[1] * isin;
            r->p0.x = S[0] + dpx * cpb * D[2] - dpy * spb;
            r->p0.y = S[1] + dpx * spb * D[2] + dpy * cpb;
            r->p0.z = S[2] - dpx * sin_t;
            r->vec.x = bx * cpb * D[2] - by * spb + bz * cpb * sin_t;
            r->vec.y = bx * spb * D[2] + by * cpb + bz * spb * sin_t;
            r->vec.z = -bx * sin_t + bz * D[2];
        } else {
            r->p0.x = S[0] + dpx;
            r->p0.y = S[1] + dpy;
            r->p0.z = S[2];
            r->vec.x = bx;
            r->vec.y = by;
            r->vec.z = (D[2] > 0.f) ? bz : -bz;
        }
        float nv = 1.f / sqrtf(r->vec.x * r->vec.x + r->vec.y * r->vec.y + r->vec.z * r->vec.z);
        r->vec.x *= nv;
        r->vec.y *= nv;
        r->vec.z *= nv;
        canfocus = 0;
    } else if (st == ORC_ST_LINE || st == ORC_ST_SLIT) {
        float t = orc_rand_next(ran);
        r->p0.x += t * P1[0];
        r->p0.y += t * P1[1];
        r->p0.z += t * P1[2];
        if (st == ORC_ST_LINE) {
            float tt = 1.f - 2.f * orc_rand_next(ran);
            float ss = 1.f - 2.f * orc_rand_next(ran);
            float q = 1.f - r->vec.x * r->vec.x - r->vec.y * r->vec.y;
            float pp = sqrtf(q > 0.f ? q : 0.f) * (orc_rand_next(ran) > 0.5f ? 1.f : -1.f);
            /* the cross product as it is, not normalized (src/mmc_raytrace.c:2552-2562) */
            v3 vv = {r->vec.y * pp - r->vec.z * ss, r->vec.z * tt - r->vec.x * pp, r->vec.x * ss - r->vec.y * tt};
            r->vec.x = vv.x;
            r->vec.y = vv.y;
            r->vec.z = vv.z;
        }
        origin.x += P1[0] * 0.5f;
        origin.y += P1[1] * 0.5f;
        origin.z += P1[2] * 0.5f;
        canfocus = (st == ORC_ST_SLIT);
    }
    if (canfocus && focus != 0.f) {     /* converging or diverging beam */
        origin.x += focus * r->vec.x;
        origin.y += focus * r->vec.y;
        origin.z += focus * r->vec.z;
        if (focus < 0.f) {
            r->vec.x = r->p0.x - origin.x;
            r->vec.y = r->p0.y - origin.y;
            r->vec.z = r->p0.z - origin.z;
        } else {
            r->vec.x = origin.x - r->p0.x;
            r->vec.y = origin.y - r->p0.y;
            r->vec.z = origin.z - r->p0.z;
        }
        float rn2 = 1.f / sqrtf(r->vec.x * r->vec.x + r->vec.y * r->vec.y + r->vec.z * r->vec.z);
        r->vec.x = r->vec.x * rn2;
        r->vec.y = r->vec.y * rn2;
        r->vec.z = r->vec.z * rn2;
    }
    r->p0 = mult_add(r->p0, r->vec, 1.f, EPS);

    for (int is = -1; is < m->srcelemlen; is++) {
        const int* elems;
        if (is < 0) {
            if (r->eid >= 0) elems = m->elem + 4 * (size_t)(r->eid - 1);
            else continue;
        } else {
            elems = m->elem + 4 * (size_t)(m->srcelem[is] - 1);
        }
        float bary[4];
        int include = 1;
        for (int i = 0; i < 4; i++) {
            v3 A = node_at(m, elems[O_out[i][0]]), B = node_at(m, elems[O_out[i][1]]),
               C = node_at(m, elems[O_out[i][2]]);
            v3 AB = diff3(A, B), AC = diff3(A, C), S = diff3(A, r->p0);
            v3 N = cross3(AB, AC);
            bary[O_facemap[i]] = -dot3(S, N);
        }
        for (int i = 0; i < 4; i++)
            if (bary[i] < -1e-4f) include = 0;
        if (include) {
            r->eid = (is >= 0 ? m->srcelem[is] : r->eid);
            float s = 0.f;
            for (int i = 0; i < 4; i++) s += bary[i];
            for (int i = 0; i < 4; i++) r->bary0[i] = bary[i] / s;
            for (int i = 0; i < 4; i++)
                if ((bary[i] / s) < 1e-4f) r->faceid = O_ifacemap[i] + 1;
            return 0;
        }
    }
    return 1;
}

typedef struct {
    float* buf;       /* reclen floats per detected photon */
    int64_t* pid;
    int n, cap, reclen;
} odet;

/* onephoton, src/mmc_raytrace.c:1772-2126 (mmMCX, no iMMC, no debug
 * output).  Returns the termination reason. */
static int orc_onephoton(int64_t id, octx* c, uint64_t ran[2], odet* det, double* launchw,
                         double* absorbw, orc_photon* rec) {
    const orc_cfg* cfg = c->cfg;
    const orc_mesh* m = c->m;
    int prop = m->prop;
    int fixcount = 0, exitdet = 0, oldeid = 0, reason = 0;
    int reclen = (2 + (cfg->ismomentum > 0)) * prop + (cfg->issaveexit > 0) * 6 + 2;
    float ppath[4096];
    double raytet0 = c->raytet;
    float rgain = 0.f, nscat = 0.f;

    oray r;
    memset(&r, 0, sizeof(r));
    r.p0.x = cfg->srcpos[0]; r.p0.y = cfg->srcpos[1]; r.p0.z = cfg->srcpos[2];
    r.vec.x = cfg->srcdir[0]; r.vec.y = cfg->srcdir[1]; r.vec.z = cfg->srcdir[2];
    memcpy(r.bary0, cfg->bary0, sizeof(float) * 4);
    r.eid = cfg->e0;
    r.faceid = -1;
    r.weight = 1.f;
    r.oldidx = 0xFFFFFFFFu;
    r.pout_undef = 1;
    memset(ppath, 0, sizeof(float) * (size_t)(reclen - 1));
    r.ppath = ppath;

    if (orc_launch(&r, c, ran)) return -1;
    float w0 = r.weight;
    if (cfg->srcnum > 1) {
        /* shared patterns record the launch pixel in the weight column
         * (src/mmc_raytrace.c:1824) and tally each pattern's launched energy */
        int32_t px = (int32_t)r.posidx;
        memcpy(&ppath[reclen - 2], &px, 4);
        if (c->srcweight)
            for (int k = 0; k < cfg->srcnum; k++) {
#pragma omp atomic
                c->srcweight[k] += cfg->srcpattern[(size_t)r.posidx * cfg->srcnum + k];
            }
    } else {
        ppath[reclen - 2] = r.weight;
    }
    const int rf = cfg->omega > 0.f && !cfg->replayweight;
    r.wmag = r.weight;
    /* replay with weighted path / scattering outputs: the detected weight
     * is the launched energy (src/mmc_raytrace.c:1815-1817) */
    *launchw += (cfg->outputtype == ORC_OT_WL || cfg->outputtype == ORC_OT_WP) ? c->rw : r.weight;

/* partial paths (src/mmc_raytrace.c:1869, 1954, 1969): column prop-1+type;
 * for a wide-field detector element (type prop+1) that is column 2 prop,
 * left where it lands as the reference does (the first momentum column, an
 * exit column rewritten at the exit, or the launch-weight column) */
#define PPATH_ADD(cond)                                                            \
    do {                                                                          \
        if (cfg->issavedet && (cond))                                              \
            ppath[prop - 1 + m->type[r.eid - 1]] += r.Lmove;                        \
    } while (0)

    while (1) {
        if (c->raytet - raytet0 >= ORC_MAX_STEPS) {
            reason = ORC_END_STEPLIMIT;
            break;
        }
        r.slen = orc_trace(&r, c);
        if (r.pout_undef) {
            if (fixcount++ < MAX_TRIAL) {
                orc_fixphoton(&r, m);
                c->nudges++;
                continue;
            }
            r.eid = ORC_ID_UNDEFINED;
            r.faceid = -1;
        }
        /* the reference reads type[eid-1] before testing faceid>=0; with
         * eid=ID_UNDEFINED faceid is -1, so no path is added */
        if (cfg->issavedet && r.Lmove > 0.f && r.faceid >= 0 && m->type[r.eid - 1] > 0)
            ppath[prop - 1 + m->type[r.eid - 1]] += r.Lmove;
        if (r.faceid == -2) break;

        while (r.faceid >= 0 && !r.isend) {
            r.p0 = r.pout;
            oldeid = r.eid;
            r.eid = m->facenb[4 * (size_t)(r.eid - 1) + r.faceid];
            float nold = med_of(m, m->type[oldeid - 1])[3];
            if (cfg->isreflect && (r.eid <= 0 || med_of(m, m->type[r.eid - 1])[3] != nold)) {
                if (!(r.eid <= 0 && ((nold == cfg->nout && cfg->isreflect != ORC_BC_MIRROR) ||
                                     cfg->isreflect == ORC_BC_ABSORBEXT))) {
                    float dir[3] = {r.vec.x, r.vec.y, r.vec.z};
                    orc_reflect(m, cfg->method, cfg->nout, cfg->isreflect, cfg->isspecular, dir,
                                oldeid, &r.eid, r.faceid, ran);
                    r.vec.x = dir[0]; r.vec.y = dir[1]; r.vec.z = dir[2];
                }
            }
            if (r.eid <= 0) break;
            if (m->type[oldeid - 1] == 0 && m->type[r.eid - 1]) {
                if (!cfg->voidtime) r.photontimer = 0.f;
            }
            /* leaving the tissue into the background ends the photon unless
             * wide-field detectors wait outside (src/mmc_raytrace.c:1930-1942) */
            if (m->type[oldeid - 1] && m->type[r.eid - 1] == 0 && !m->isextdet) {
                r.eid = 0;
                break;
            }
            r.slen = orc_trace(&r, c);
            PPATH_ADD(r.Lmove > 0.f && m->type[r.eid - 1] > 0);
            fixcount = 0;
            while (r.pout_undef && fixcount++ < MAX_TRIAL) {
                orc_fixphoton(&r, m);
                c->nudges++;
                r.slen = orc_trace(&r, c);
                PPATH_ADD(r.Lmove > 0.f && m->type[r.eid - 1] > 0);
            }
            if (r.pout_undef) {
                r.eid = ORC_ID_UNDEFINED;
                break;
            }
        }

        if (r.eid <= 0 || r.pout_undef) {
            if (r.eid != ORC_ID_UNDEFINED && cfg->issavedet && cfg->issaveexit) {
                ppath[reclen - 2 - 6] = r.p0.x;
                ppath[reclen - 2 - 5] = r.p0.y;
                ppath[reclen - 2 - 4] = r.p0.z;
                ppath[reclen - 2 - 3] = r.vec.x;
                ppath[reclen - 2 - 2] = r.vec.y;
                ppath[reclen - 2 - 1] = r.vec.z;
            }
            if (cfg->issavedet && r.eid <= 0 && m->isextdet && cfg->detnum == 0 &&
                m->type[oldeid - 1] == prop + 1) {
                /* left the mesh through a detector element: its id is the
                 * detector id (src/mmc_raytrace.c:2013-2014) */
                exitdet = oldeid;
            } else if (cfg->issavedet && r.eid <= 0) {
                for (int i = 0; i < cfg->detnum; i++) {
                    const float* d = cfg->detpos + 4 * i;
                    if ((d[0] - r.p0.x) * (d[0] - r.p0.x) + (d[1] - r.p0.y) * (d[1] - r.p0.y) +
                            (d[2] - r.p0.z) * (d[2] - r.p0.z) <
                        d[3] * d[3]) {
                        exitdet = i + 1;
                        break;
                    }
                }
            }
            reason = (r.eid == ORC_ID_UNDEFINED) ? ORC_END_LOST : ORC_END_EXIT;
            break;
        }

        /* RF: the roulette tests the magnitude of the complex weight and
         * scales both parts (src/mmc_core.cl:2090-2110) */
        if (cfg->minenergy > 0.f && (rf ? r.wmag : r.weight) < cfg->minenergy &&
            (cfg->tend - cfg->tstart) * c->rtstep <= 1.f) {
            if (orc_rand_next(ran) * cfg->roulettesize <= 1.f) {
                float wb = rf ? r.wmag : r.weight;
                r.weight *= cfg->roulettesize;
                if (rf) {
                    r.weight_im *= cfg->roulettesize;
                    r.wmag *= cfg->roulettesize;
                }
                rgain += (rf ? r.wmag : r.weight) - wb;
            } else {
                reason = ORC_END_ROULETTE;
                break;
            }
        }

        {
            float dir[3] = {r.vec.x, r.vec.y, r.vec.z};
            float mom = 0.f;
            r.slen0 = orc_scatter_mom(med_of(m, m->type[r.eid - 1])[2], dir, ran, &mom);
            r.vec.x = dir[0]; r.vec.y = dir[1]; r.vec.z = dir[2];
            if (cfg->ismomentum && cfg->issavedet && m->type[r.eid - 1] > 0)
                ppath[2 * prop - 1 + m->type[r.eid - 1]] += mom;   /* third block: momentum transfer */
        }
        r.slen = r.slen0;
        if (m->type[r.eid - 1] > 0) ppath[m->type[r.eid - 1] - 1]++;
        nscat += 1.f;
    }
#undef PPATH_ADD

    if (cfg->issavedet && exitdet > 0 && det && det->buf) {
        int slot;
#pragma omp atomic capture
        slot = det->n++;
        if (slot < det->cap) {
            float* row = det->buf + (size_t)slot * reclen;
            row[0] = (float)exitdet;
            memcpy(row + 1, ppath, sizeof(float) * (size_t)(reclen - 1));
            if (det->pid) det->pid[slot] = id;
        }
    }
    *absorbw += r.Eabsorb;
    if (cfg->srcnum > 1 && c->srcweight)
        for (int k = 0; k < cfg->srcnum; k++) {
#pragma omp atomic
            c->srcweight[cfg->srcnum + k] += (double)r.Eabsorb * cfg->srcpattern[(size_t)r.posidx * cfg->srcnum + k];
        }

    if (rec) {
        rec->p[0] = r.p0.x; rec->p[1] = r.p0.y; rec->p[2] = r.p0.z;
        rec->v[0] = r.vec.x; rec->v[1] = r.vec.y; rec->v[2] = r.vec.z;
        rec->w = (reason == ORC_END_ROULETTE) ? 0.f : (rf ? r.wmag : r.weight);
        rec->t = r.photontimer;
        rec->eid = r.eid;
        rec->reason = reason;
        rec->nraytet = (int32_t)(c->raytet - raytet0);
        rec->detid = exitdet;
        rec->eabs = (float)r.Eabsorb;
        rec->rgain = (reason == ORC_END_ROULETTE) ? rgain - (rf ? r.wmag : r.weight) : rgain;
        rec->w0 = w0;
        rec->nscat = nscat;
    }
    return reason;
}

static double now_ms(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

/* mmc_run_mp, src/mmc_host.c:178-468: OpenMP loop over photons.  With
 * cfg->photonseed each photon runs from its own saved xorshift128+ state (the
 * reference's seed-replay mode, src/mmc_host.c:280-281); otherwise each thread
 * is seeded from srand(seed); rand() words (src/mmc_host.c:238-246). */
int orc_run(const orc_cfg* cfg, const orc_mesh* mesh, orc_out* out) {
    octx base;
    base.cfg = cfg;
    base.m = mesh;
    base.weight = out->weight;
    base.weight_im = out->weight_im;
    base.srcweight = out->srcweight;
    base.rtstep = 1.f / cfg->tstep;
    base.framelen = (cfg->method == ORC_RT_BLBADOUELGRID)
                        ? cfg->griddim[0] * cfg->griddim[1] * cfg->griddim[2]
                        : (cfg->basisorder ? mesh->nn : mesh->ne);
    base.raytet = 0.0;
    int reclen = (2 + (cfg->ismomentum > 0)) * mesh->prop + (cfg->issaveexit > 0) * 6 + 2;
    odet det = {out->detdata, out->detphoton, 0, out->maxdet, reclen};
    double launch = 0.0, absorb = 0.0, raytet = 0.0;
    int err = 0;
    double t0 = now_ms();
    int nth = cfg->nthreads;
#ifdef _OPENMP
    if (nth <= 0) nth = omp_get_max_threads();
#else
    nth = 1;
#endif
    uint32_t* seeds = NULL;
    if (!cfg->photonseed) {
        seeds = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (size_t)nth);
        srand(cfg->seed);
        for (int i = 0; i < 4 * nth; i++) seeds[i] = (uint32_t)rand();
    }

#pragma omp parallel num_threads(nth) reduction(+ : launch, absorb, raytet) reduction(| : err)
    {
        octx c = base;
        uint64_t tran[2] = {0, 0};
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        if (seeds) orc_rand_seed(seeds + 4 * tid, tran);
        /* Kahan sums per thread as in src/mmc_raytrace.c:1812-1838 */
        double lsum = 0.0, asum = 0.0;
#pragma omp for schedule(dynamic, 64)
        for (int64_t id = 0; id < cfg->nphoton; id++) {
            uint64_t pran[2];
            uint64_t* ran = tran;
            if (cfg->photonseed) {
                pran[0] = cfg->photonseed[2 * id];
                pran[1] = cfg->photonseed[2 * id + 1];
                ran = pran;
            }
            orc_photon* rec = NULL;
            if (out->rec && id >= cfg->recfirst && id < cfg->recfirst + cfg->reccount)
                rec = out->rec + (id - cfg->recfirst);
            double lw = 0.0, aw = 0.0;
            c.rw = cfg->replayweight ? cfg->replayweight[id] : 1.f;
            c.rt = cfg->replaytime ? cfg->replaytime[id] : 0.f;
            c.ties = 0;
            c.nudges = 0;
            if (orc_onephoton(id, &c, ran, &det, &lw, &aw, rec) < 0) err = 1;
            if (rec && out->ties) {
                out->ties[2 * (id - cfg->recfirst)] = c.ties;
                out->ties[2 * (id - cfg->recfirst) + 1] = c.nudges;
            }
            lsum += lw;
            asum += aw;
        }
        launch += lsum;
        absorb += asum;
        raytet += c.raytet;
    }
    free(seeds);
    out->launchweight += launch;
    out->absorbweight += absorb;
    out->raytet += raytet;
    out->detected = det.n;
    out->reclen = reclen;
    out->runtime_ms = now_ms() - t0;
    out->error = err;
    return err ? -1 : 0;
}

/* mesh_normalize, src/mmc_mesh.c:2154-2279 (single source, no RF). */
double orc_normalize(const orc_cfg* cfg, const orc_mesh* mesh, double* weight, float Eabsorb,
                     float Etotal) {
    const int grid = cfg->method == ORC_RT_BLBADOUELGRID;
    int datalen = grid ? cfg->griddim[0] * cfg->griddim[1] * cfg->griddim[2]
                       : (cfg->basisorder ? mesh->nn : mesh->ne);
    double energydeposit = 0.f, normalizor;
    if (cfg->photonseed && (cfg->outputtype == ORC_OT_JACOBIAN || cfg->outputtype == ORC_OT_WL ||
                            cfg->outputtype == ORC_OT_WP)) {
        float nz = 1.f / (DELTA_MUA * cfg->nphoton);
        if (cfg->outputtype == ORC_OT_WL || cfg->outputtype == ORC_OT_WP) nz = 1.f / Etotal;
        for (size_t i = 0; i < (size_t)cfg->maxgate * datalen; i++) weight[i] *= nz;
        return nz;
    }
    if (cfg->outputtype == ORC_OT_ENERGY) {
        normalizor = 1.f / Etotal;
        for (size_t i = 0; i < (size_t)cfg->maxgate * datalen; i++) weight[i] *= normalizor;
        return normalizor;
    }
    if (grid) {
        /* voxel fluence per unit launched energy (src/mmc_mesh.c:2205-2206) */
        normalizor = 1.0 / (Etotal * cfg->unitinmm * cfg->unitinmm * cfg->unitinmm);
    } else if (cfg->basisorder) {
        for (int i = 0; i < cfg->maxgate; i++)
            for (int j = 0; j < datalen; j++)
                if (mesh->nvol[j] > 0.f) weight[(size_t)i * datalen + j] /= mesh->nvol[j];
        for (int i = 0; i < mesh->ne; i++) {
            const int* ee = mesh->elem + 4 * (size_t)i;
            double energyelem = 0.f;
            for (int j = 0; j < cfg->maxgate; j++)
                for (int k = 0; k < 4; k++) {
                    float re_val = (float)weight[(size_t)j * mesh->nn + ee[k] - 1];
                    energyelem += re_val;
                }
            energydeposit += energyelem * mesh->evol[i] * med_of(mesh, mesh->type[i])[0];
        }
        normalizor = Eabsorb / (Etotal * energydeposit * 0.25f);
    } else {
        for (int i = 0; i < datalen; i++)
            for (int j = 0; j < cfg->maxgate; j++) energydeposit += weight[(size_t)j * datalen + i];
        for (int i = 0; i < datalen; i++) {
            double energyelem = mesh->evol[i] * med_of(mesh, mesh->type[i])[0];
            for (int j = 0; j < cfg->maxgate; j++) weight[(size_t)j * datalen + i] /= energyelem;
        }
        normalizor = Eabsorb / (Etotal * energydeposit);
    }
    if (cfg->outputtype == ORC_OT_FLUX) normalizor /= cfg->tstep;
    for (size_t i = 0; i < (size_t)cfg->maxgate * datalen; i++) weight[i] *= normalizor;
    return normalizor;
}
