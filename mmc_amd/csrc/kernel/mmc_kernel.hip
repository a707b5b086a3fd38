/*
 * mmc_kernel.hip -- persistent photon-transport kernel for gfx950 (MI355X).
 *
 * One photon per lane.  Every iteration of the persistent loop performs one
 * ray-tetrahedron test per live lane followed by the bookkeeping the
 * reference's onephoton() does after that test (src/mmc_raytrace.c:1772-2126):
 * face crossing, Fresnel reflection, exit/detector capture, roulette and
 * scattering.  The nested loops of onephoton are flattened into a per-lane
 * state machine (ctx = OUTER / INNER / INNERFIX) so that every lane of a
 * wavefront runs its tracer in the same iteration.  Lanes whose photon ended
 * are refilled from a device work queue: a wave ballot counts the dead lanes,
 * one lane claims that many photon ids with a single atomic, and each dead
 * lane takes base + (number of dead lanes below it).
 *
 * Mesh arrays are structure-of-arrays in HBM.  Per-tissue optical properties
 * and per-simulation parameters are staged in LDS; per-lane energy tallies and
 * partial-path buffers live in LDS too.  Fluence deposits are coalesced in a
 * register while the photon stays in one element and time gate (as
 * src/mmc_raytrace.c:1562-1612 does) and then written with FP32 atomics.
 *
 * All per-photon arithmetic is FP32, in the reference CPU's operation order
 * (compiled with -ffp-contract=off); FP64 appears only in the per-simulation
 * energy totals written when a lane changes simulation or exits.
 */
#include <hip/hip_runtime.h>

#include "mmc_fp32math.h"
#include "mmc_kernel.h"

namespace mmc {

#define EPSF       1e-6f                      /* src/mmc_rand_common.h:46 */
#define R_C0F      3.335640951981520e-12f     /* src/mmc_const.h:45 */
#define R_MIN_MUSF 1e9f                       /* src/mmc_mesh.h:72 */
#define ID_UNDEF   0x7FFFFFFF                 /* src/mmc_mesh.h:66 */
#define MAX_TRIAL  3                          /* src/mmc_raytrace.h:44 */
#define FIX_PHOTON 1e-3f                      /* src/mmc_raytrace.h:45 */
#define DELTA_MUA  1e-4f                      /* src/mmc_mesh.h:73 */

enum { CTX_OUTER = 0, CTX_INNER = 1, CTX_INNERFIX = 2 };

/* The face tables of src/mmc_raytrace.c:75-152 and src/mmc_mesh.c:59-103
 * appear below as selects and packed nibbles so that every index resolves in
 * registers:
 *   Plucker face -> edge products  fc = {0,4,2} {3,5,4} {2,5,1} {1,3,0}
 *   Plucker face -> nodes          nc = {3,0,1} {3,1,2} {2,0,3} {1,0,2}
 *   launch sub-volume faces        out = {0,3,1} {3,2,1} {0,2,3} {0,1,2},
 *                                  written to barycentric slot {2,0,1,3} */

/* plane index -> facenb column, and its inverse */
__device__ __forceinline__ int faceorder(int i) { return (0x1320 >> (4 * (3 - i))) & 0xF; }
__device__ __forceinline__ int ifaceorder(int f) { return (0x3021 >> (4 * (3 - f))) & 0xF; }

__device__ __forceinline__ int pick4(int4 v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
__device__ __forceinline__ float pick4f(float4 v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
__device__ __forceinline__ float pick4a(const float* v, int i) {
    return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3]));
}
/* entry i (0..3) of a table packed as four nibbles, most significant first */
__device__ __forceinline__ int nib4(int packed, int i) { return (packed >> (4 * (3 - i))) & 0xF; }
/* face neighbour: (element << 4) | media label of that element (lblshift
 * 4), or the element id alone (lblshift 0; the label is read from type) */
__device__ __forceinline__ int nb_id(int packed, int lblshift) { return packed >> lblshift; }

/* ---- xorshift128+ (src/mmc_rand_xorshift128p.c:55-70) ------------------- */
__device__ __forceinline__ float rng_next(uint64_t& s0r, uint64_t& s1r) {
    uint64_t s1 = s0r;
    const uint64_t s0 = s1r;
    s0r = s0;
    s1 ^= s1 << 23;
    s1r = s1 ^ s0 ^ (s1 >> 18) ^ (s0 >> 5);
    uint32_t lo = (uint32_t)(s1r + s0);
    return __uint_as_float(0x3F800000u | (lo >> 9)) - 1.0f;
}
/* the generator output k (23 bits) behind a draw u = k / 2^23 */
__device__ __forceinline__ uint32_t draw_bits(float u) { return __float_as_uint(u + 1.0f) & 0x7FFFFFu; }

/* ---- host C library agreement (built by mmc_host_libm_table) ------------
 * The reference draws scattering lengths, azimuths and isotropic zeniths
 * through the host C library's logf/sincosf/acosf.  Each of those calls sees
 * a function of one generator output k.  The kernel evaluates the value in
 * double and rounds it to float f; the C library returns either f or the
 * float o next to f on the far side of the double, and a host-built bit per
 * k and value says which.  The bit can only be set where the double lies
 * near the rounding midpoint of f and o: within D |o - f| of it, D being the
 * threshold of the region of k (2^MMC_LIBM_REGSHIFT consecutive outputs), so
 * the bits are read for a few percent of the draws only.  One copy per
 * device, filled once per process. */
__device__ uint8_t g_lm_log[MMC_FIXB_LOG];      /* logf(u + 1e-6f): 1 bit per k */
__device__ uint8_t g_lm_az[MMC_FIXB_AZ];        /* sincosf((float)(2 pi u)): sin, cos */
__device__ uint8_t g_lm_iso[MMC_FIXB_ISO];      /* acosf(2u - 1), then sincosf of it: 4 bits */
__device__ uint8_t g_lm_pi[MMC_FIXB_PI];        /* sincosf((float)(pi u)): sin, cos */
constexpr int kThrLog = 0, kThrAz = MMC_LIBM_NREG, kThrIso = 3 * MMC_LIBM_NREG, kThrPi = 6 * MMC_LIBM_NREG;
__device__ float g_lm_thr[8 * MMC_LIBM_NREG];    /* [table][region][value] thresholds */
/* coefficients of the double evaluations (mmc_fp32math.h), in the constant
 * address space so that wave-uniform reads of them are scalar loads */
__constant__ double g_dcoef[MMC_NDCOEF] = {MMC_DCOEF_LIST};
typedef const __attribute__((address_space(4))) double* kconst_ptr;

/* the coefficient table, through an opaque copy of its address: its entries
 * are loaded (scalar loads) where a polynomial needs them instead of being
 * held in registers across the photon loop */
__device__ __forceinline__ const double* dk() {
    kconst_ptr p = (kconst_ptr)g_dcoef;
    asm volatile("" : "+s"(p));
    return (const double*)p;
}


/* The launch arguments, through an opaque copy of the kernel-argument
 * segment's address (the by-value KArgs parameter sits at its start).  The
 * transport loop takes a fresh copy every iteration, so the arguments are
 * reloaded (scalar loads) within the iteration that uses them instead of
 * being held in scalar registers across the loop: held, they overflowed the
 * scalar register file into lanes of vector registers (147 spilled scalar
 * registers, restored by v_readlane in the loop). */
typedef const __attribute__((address_space(4))) KArgs* kargs_ptr;
__device__ __forceinline__ const KArgs* kargs() {
    kargs_ptr p = (kargs_ptr)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return (const KArgs*)p;
}

/* true when the C library may round e to the far neighbour o of f =
 * (float)e: e within D |o - f| of the midpoint.  With r = e - f and u = |o -
 * f| (one ulp of f, half of it when stepping towards zero from a power of
 * two; a power of two either way) that is |r| >= (0.5 - D) u, and every
 * operation is exact.  (Checked against the midpoint form on all 2^23
 * generator outputs of every draw and thresholds on the boundary.) */
__device__ __forceinline__ bool lm_near(double e, float f, float D) {
    const double r = e - (double)f;
    if (r == 0.0) return false;
    const uint32_t b = __float_as_uint(f) & 0x7FFFFFFFu;
    int ex = (int)(b >> 23);
    const bool away = (r > 0.0) == (f > 0.f);
    const int half = (!away && (b & 0x7FFFFFu) == 0u && ex > 1) ? 1 : 0;
    ex = ex > 1 ? ex : 1;
    const double u = __longlong_as_double((long long)(ex - 150 - half + 1023) << 52);
    return fabs(r) >= (0.5 - (double)D) * u;
}
__device__ __forceinline__ float lm_pick(float f, double e, uint32_t bit) { return bit ? mmc_fnext(f, e) : f; }

/* The thresholds of the two draws of every scattering event (log, then
 * azimuth sin and cos) are staged in LDS by each workgroup: thr_lds =
 * [log: MMC_LIBM_NREG][azimuth: MMC_LIBM_NREG x 2]. */
constexpr int kThrLds = 3 * MMC_LIBM_NREG;

/* -logf(u + EPS) of a draw u as the host C library rounds it
 * (rand_next_scatlen, src/mmc_rand_common.h:50-51); tab: the table byte of
 * the draw's generator output (g_lm_log[k >> 3]), fetched by the caller */
__device__ __forceinline__ float scatlen_of(float u, uint32_t tab, const float* thr_lds) {
    const uint32_t k = draw_bits(u);
    const double e = mmc_log_dk(u + EPSF, dk());
    float l = (float)e;
    if (lm_near(e, l, thr_lds[k >> MMC_LIBM_REGSHIFT])) l = lm_pick(l, e, (tab >> (k & 7)) & 1u);
    return -l;
}
__device__ __forceinline__ float rng_scatlen(uint64_t& a, uint64_t& b, const float* thr_lds) {
    const float u = rng_next(a, b);
    return scatlen_of(u, g_lm_log[draw_bits(u) >> 3], thr_lds);
}
/* acosf(2u - 1) of a draw u, as the host C library rounds it */
__device__ __forceinline__ float acos_corrected(float u) {
    const uint32_t k = draw_bits(u);
    const double et = mmc_acos_dk(2.f * u - 1.f, dk());
    float t = (float)et;
    if (lm_near(et, t, g_lm_thr[kThrIso + 3 * (k >> MMC_LIBM_REGSHIFT)]))
        t = lm_pick(t, et, (g_lm_iso[k >> 1] >> (4 * (k & 1))) & 1u);
    return t;
}

/* kinds of angle evaluated by angle_draw */
enum { kAngNone = 0, kAngAz = 1, kAngPi = 2, kAngIso = 3, kAngCr = 4 };

/* sin and cos of one angle: kAngAz (float)(2 pi u), kAngPi (float)(pi u) or
 * kAngIso acosf(2u - 1) of a draw u, each as the host C library rounds it
 * (the isotropic zenith's sine and cosine are those of the C library's
 * angle), or kAngCr an angle given directly (|arg| < 2^20; the double
 * evaluation rounded, i.e. correctly rounded except within about 2^-29 ulp of
 * a midpoint, where the C library's sinf/cosf, within 0.56 ulp, agree
 * wherever they round correctly).  Callers keep one call site per path, so
 * the double evaluation is emitted once there. */
__device__ __forceinline__ void angle_draw(int kind, float arg, float* s, float* c, const float* thr_lds) {
    const uint32_t k = draw_bits(arg);
    const float t = kind == kAngAz    ? mmc_twopi(arg)
                    : kind == kAngPi  ? mmc_pi(arg)
                    : kind == kAngIso ? acos_corrected(arg)
                                      : arg;
    double es, ec;
    mmc_sincos_dk(t, &es, &ec, dk());
    float fs = (float)es, fc = (float)ec;
    if (kind != kAngCr) {
        const uint32_t reg = k >> MMC_LIBM_REGSHIFT;
        const float* thr = kind == kAngAz ? thr_lds + MMC_LIBM_NREG + 2 * reg
                                          : g_lm_thr + (kind == kAngPi ? kThrPi + 2 * reg : kThrIso + 3 * reg + 1);
        if (lm_near(es, fs, thr[0]) || lm_near(ec, fc, thr[1])) {
            const uint32_t code = kind == kAngIso ? (g_lm_iso[k >> 1] >> (4 * (k & 1) + 1))
                                                  : ((kind == kAngAz ? g_lm_az : g_lm_pi)[k >> 2] >> (2 * (k & 3)));
            fs = lm_pick(fs, es, code & 1u);
            fc = lm_pick(fc, ec, code & 2u);
        }
    }
    *s = fs;
    *c = fc;
}

/* Copies the host-built agreement tables into this device's copy. */
int upload_libm_tables(const uint8_t* blog, const uint8_t* baz, const uint8_t* biso, const uint8_t* bpi,
                       const float* tlog, const float* taz, const float* tiso, const float* tpi) {
    const size_t nr = MMC_LIBM_NREG * sizeof(float);
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_lm_log), blog, MMC_FIXB_LOG) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(g_lm_az), baz, MMC_FIXB_AZ) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(g_lm_iso), biso, MMC_FIXB_ISO) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(g_lm_pi), bpi, MMC_FIXB_PI) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(g_lm_thr), tlog, nr, kThrLog * sizeof(float)) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(g_lm_thr), taz, 2 * nr, kThrAz * sizeof(float)) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(g_lm_thr), tiso, 3 * nr, kThrIso * sizeof(float)) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(g_lm_thr), tpi, 2 * nr, kThrPi * sizeof(float)) != hipSuccess)
        return report_error(-1, "libm agreement: table upload failed");
    return 0;
}

/* rotatevector (src/mmc_mesh.h:336-357): turn (vx, vy, vz) by the zenith
 * (st, ct) and azimuth (sp, cp), then renormalize */
__device__ __forceinline__ void rotate_dir(float& vx, float& vy, float& vz, float st, float ct, float sp, float cp) {
    float qx, qy, qz;
    if (vz > -1.f + EPSF && vz < 1.f - EPSF) {
        const float t0 = 1.f - vz * vz;
        const float t1 = st * (1.f / sqrtf(t0));
        qx = t1 * (vx * vz * cp - vy * sp) + vx * ct;
        qy = t1 * (vy * vz * cp + vx * sp) + vy * ct;
        qz = -t1 * t0 * cp + vz * ct;
    } else {
        qx = st * cp;
        qy = st * sp;
        qz = (vz > 0.f) ? ct : -ct;
    }
    const float n = 1.f / sqrtf(qx * qx + qy * qy + qz * qz);
    vx = qx * n;
    vy = qy * n;
    vz = qz * n;
}

constexpr float kJustBelowOne = 0.9998f;   /* pattern index clamp, src/mmc_raytrace.c:54 */
constexpr int kConeTries = 1 << 16;        /* bound of the cone source's rejection loop */

/* launchphoton (src/mmc_raytrace.c:2346-2592): start position, direction and
 * weight of a photon for every source type, from the photon's stream after
 * its first scattering length.  Returns true when the start moved off the
 * source point, so the enclosing element has to be searched (after the EPS
 * nudge along the direction); pencil, cone, isotropic and arcsine sources keep
 * the given element and barycentric weights.  posidx: the pattern pixel. */
__device__ __forceinline__ bool launch_source(const SimCold& C, const float* pattern, uint64_t& r0, uint64_t& r1,
                                              float& px, float& py, float& pz, float& vx, float& vy, float& vz,
                                              float& w, int& posidx, const float* thr_lds) {
    const int st = C.srctype;
    const float4 P1 = C.srcparam1, P2 = C.srcparam2, S = C.srcpos, D = C.srcdir;
    float ox = S.x, oy = S.y, oz = S.z;      /* focus origin */
    bool canfocus = true;
    if (st == MMC_SRC_PENCIL) return false;
    const bool planar = st == MMC_SRC_PLANAR || st == MMC_SRC_PATTERN || st == MMC_SRC_PATTERN3D ||
                        st == MMC_SRC_FOURIER || st == MMC_SRC_PENCILARRAY;
    const bool fourierx = st == MMC_SRC_FOURIERX || st == MMC_SRC_FOURIERX2D;
    const bool disk = st == MMC_SRC_DISK || st == MMC_SRC_RING || st == MMC_SRC_GAUSSIAN;
    const bool dirsrc = st == MMC_SRC_CONE || st == MMC_SRC_ISOTROPIC || st == MMC_SRC_ARCSINE;
    /* Every source needs at most two angles; each is requested as (kind,
     * argument) in the order the reference draws it and evaluated by one
     * angle_draw, so that the double evaluation is emitted twice here. */
    int kind = kAngNone;
    float arg = 0.f;
    /* ---- phase 1: draws up to the first angle ---- */
    float rx = 0.f, ry = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
    if (planar) {
        rx = rng_next(r0, r1);
        ry = rng_next(r0, r1);
        if (st == MMC_SRC_PATTERN3D) {
            const float rz = rng_next(r0, r1);
            px = S.x + rx * P1.x;
            py = S.y + ry * P1.y;
            pz = S.z + rz * P1.z;
            const int nx = (int)P1.x, ny = (int)P1.y, nz = (int)P1.z;
            const int iz = min((int)(rz * kJustBelowOne * (float)nz), nz - 1);
            const int iy = min((int)(ry * kJustBelowOne * (float)ny), ny - 1);
            const int ix = min((int)(rx * kJustBelowOne * (float)nx), nx - 1);
            w = (C.srcnum > 1) ? 1.f : pattern[C.patbase + (iz * ny + iy) * nx + ix];
            ox += P1.x * 0.5f;
            oy += P1.y * 0.5f;
            oz += P1.z * 0.5f;
        } else if (st == MMC_SRC_PENCILARRAY) {
            const float fx = floorf(rx * P1.w), fy = floorf(ry * P2.w);
            px = S.x + fx * P1.x / (P1.w - 1.f) + fy * P2.x / (P2.w - 1.f);
            py = S.y + fx * P1.y / (P1.w - 1.f) + fy * P2.y / (P2.w - 1.f);
            pz = S.z + fx * P1.z / (P1.w - 1.f) + fy * P2.z / (P2.w - 1.f);
            ox += (P1.x + P2.x) * 0.5f;
            oy += (P1.y + P2.y) * 0.5f;
            oz += (P1.z + P2.z) * 0.5f;
        } else {
            px = S.x + rx * P1.x + ry * P2.x;
            py = S.y + rx * P1.y + ry * P2.y;
            pz = S.z + rx * P1.z + ry * P2.z;
            if (st == MMC_SRC_PATTERN) {
                const int xs = C.patnx, ys = C.patny;
                const int iy = min((int)(ry * ys), ys - 1);
                const int ix = min((int)(rx * xs), xs - 1);
                posidx = iy * xs + ix;
                w = (C.srcnum == 1) ? pattern[C.patbase + posidx] : 1.f;
            } else if (st == MMC_SRC_FOURIER) {
                const float kx = floorf(P1.w), ky = floorf(P2.w);
                kind = kAngCr;
                arg = mmc_twopi(kx * rx + ky * ry + P1.w - kx);
            }
            ox += (P1.x + P2.x) * 0.5f;
            oy += (P1.y + P2.y) * 0.5f;
            oz += (P1.z + P2.z) * 0.5f;
        }
    } else if (fourierx) {
        rx = rng_next(r0, r1);
        ry = rng_next(r0, r1);
        const float vw = P1.w * (1.f / sqrtf(P1.x * P1.x + P1.y * P1.y + P1.z * P1.z));
        ux = vw * (D.y * P1.z - D.z * P1.y);
        uy = vw * (D.z * P1.x - D.x * P1.z);
        uz = vw * (D.x * P1.y - D.y * P1.x);
        px = S.x + rx * P1.x + ry * ux;
        py = S.y + rx * P1.y + ry * uy;
        pz = S.z + rx * P1.z + ry * uz;
        kind = kAngCr;
        arg = (st == MMC_SRC_FOURIERX2D) ? mmc_twopi(P2.x * rx + P2.z) : mmc_twopi(P2.x * rx + P2.y * ry + P2.z);
        ox += (P1.x + ux) * 0.5f;
        oy += (P1.y + uy) * 0.5f;
        oz += (P1.z + uz) * 0.5f;
    } else if (disk && st == MMC_SRC_RING && (P1.z > 0.f || P1.w > 0.f)) {
        kind = kAngCr;                        /* ring sector: azimuth between two angles */
        arg = fabsf(P1.z - P1.w) * rng_next(r0, r1) + fminf(P1.z, P1.w);
    } else if (disk || dirsrc || st == MMC_SRC_ZGAUSSIAN || st == MMC_SRC_HYPERBOLOID) {
        kind = kAngAz;                        /* azimuth of the draw */
        arg = rng_next(r0, r1);
    }
    float s1 = 0.f, c1 = 1.f;
    if (kind != kAngNone) angle_draw(kind, arg, &s1, &c1, thr_lds);
    /* ---- phase 2: draws up to the second angle ---- */
    kind = kAngNone;
    if (st == MMC_SRC_FOURIERX2D) {
        kind = kAngCr;
        arg = mmc_twopi(P2.y * ry + P2.w);
    } else if (st == MMC_SRC_CONE) {
        /* rejection of zeniths beyond the half angle (uniform in angle or in
         * solid angle), bounded */
        int k = 0;
        float ang, uu;
        do {
            uu = rng_next(r0, r1);
            ang = (P1.y > 0.f) ? mmc_twopi(uu) : acos_corrected(uu);
        } while (ang > P1.x && ++k < kConeTries);
        kind = (ang > P1.x) ? kAngCr : ((P1.y > 0.f) ? kAngAz : kAngIso);
        arg = (ang > P1.x) ? P1.x : uu;
    } else if (st == MMC_SRC_ISOTROPIC) {
        kind = kAngIso;
        arg = rng_next(r0, r1);
    } else if (st == MMC_SRC_ARCSINE) {
        kind = kAngPi;
        arg = rng_next(r0, r1);
    } else if (st == MMC_SRC_ZGAUSSIAN) {
        /* the second draw comes from the photon's own stream: the
         * reference takes it from ran0, which in mmc_run_mp is a
         * thread-private generator state that rng_init never seeds
         * (src/mmc_host.c:227,265,280-283; src/mmc_rand_xorshift128p.c:82-84),
         * so its values cannot be reproduced */
        const float u1 = rng_next(r0, r1);
        const float u2 = rng_next(r0, r1);
        kind = kAngCr;
        arg = sqrtf((float)(-2.0 * log((double)u1))) * (1.f - 2.f * u2) * P1.x;
    }
    float s2 = 0.f, c2 = 1.f;
    if (kind != kAngNone) angle_draw(kind, arg, &s2, &c2, thr_lds);
    /* ---- phase 3: positions, weights and directions ---- */
    if (st == MMC_SRC_FOURIER) {
        w = (c1 * (1.f - P2.w + floorf(P2.w)) + 1.f) * 0.5f;
    } else if (fourierx) {
        w = (st == MMC_SRC_FOURIERX2D) ? (s1 * s2 + 1.f) * 0.5f : (c1 * (1.f - P2.w) + 1.f) * 0.5f;
    } else if (disk) {
        const float sphi = s1, cphi = c1;
        float rr;
        /* the Gaussian radius takes the double-precision log of the draw,
         * converted to float for sqrtf (src/mmc_raytrace.c:2444-2449) */
        if (st != MMC_SRC_GAUSSIAN) {
            rr = sqrtf(rng_next(r0, r1) * fabsf(P1.x * P1.x - P1.y * P1.y) + P1.y * P1.y);
        } else if (fabsf(D.w) < 1e-5f || fabsf(P1.y) < 1e-5f) {
            rr = sqrtf((float)(-log((double)rng_next(r0, r1)))) * P1.x;
        } else {
            const float z0 = (float)((double)(P1.x * P1.x) * 3.141592653589793 / (double)P1.y);  /* Rayleigh range */
            const float f = 1.f + (D.w * D.w / (z0 * z0));
            rr = sqrtf((float)(-log((double)rng_next(r0, r1)) * (double)f)) * P1.x;
        }
        if (D.z > -1.f + EPSF && D.z < 1.f - EPSF) {
            const float t0 = 1.f - D.z * D.z;
            const float t1 = rr / sqrtf(t0);
            px = S.x + t1 * (D.x * D.z * cphi - D.y * sphi);
            py = S.y + t1 * (D.y * D.z * cphi + D.x * sphi);
            pz = S.z - t1 * t0 * cphi;
        } else {
            px = S.x + rr * cphi;
            py = S.y + rr * sphi;
        }
    } else if (dirsrc) {
        rotate_dir(vx, vy, vz, s2, c2, s1, c1);
        return false;
    } else if (st == MMC_SRC_ZGAUSSIAN) {
        rotate_dir(vx, vy, vz, s2, c2, s1, c1);
        canfocus = false;
    } else if (st == MMC_SRC_HYPERBOLOID) {
        const float sphi = s1, cphi = c1;
        const float rr = sqrtf(0.5f * rng_scatlen(r0, r1, thr_lds)) * P1.x;   /* waist radius */
        const float tilt = (P1.z > 0.f) ? (-P1.y / P1.z) : 0.f;
        float bx = -rr * sphi, by = rr * cphi, bz = P1.z;
        const float nb = 1.f / sqrtf(bx * bx + by * by + bz * bz);
        bx *= nb;
        by *= nb;
        bz *= nb;
        const float dx = rr * (cphi - tilt * sphi), dy = rr * (sphi + tilt * cphi);
        if (D.z > -1.f + EPSF && D.z < 1.f - EPSF) {
            const float sth = sqrtf(1.f - D.z * D.z);
            const float ist = 1.f / sth;
            const float cpb = D.x * ist, spb = D.y * ist;
            px = S.x + dx * cpb * D.z - dy * spb;
            py = S.y + dx * spb * D.z + dy * cpb;
            pz = S.z - dx * sth;
            vx = bx * cpb * D.z - by * spb + bz * cpb * sth;
            vy = bx * spb * D.z + by * cpb + bz * spb * sth;
            vz = -bx * sth + bz * D.z;
        } else {
            px = S.x + dx;
            py = S.y + dy;
            pz = S.z;
            vx = bx;
            vy = by;
            vz = (D.z > 0.f) ? bz : -bz;
        }
        const float nv = 1.f / sqrtf(vx * vx + vy * vy + vz * vz);
        vx *= nv;
        vy *= nv;
        vz *= nv;
        canfocus = false;
    } else if (st == MMC_SRC_LINE || st == MMC_SRC_SLIT) {
        const float t = rng_next(r0, r1);
        px = S.x + t * P1.x;
        py = S.y + t * P1.y;
        pz = S.z + t * P1.z;
        if (st == MMC_SRC_LINE) {
            /* a direction across the line: the cross product as the
             * reference leaves it, not normalized (src/mmc_raytrace.c:2552-2562);
             * the max() only keeps a rounding-negative radicand from
             * producing NaN */
            const float tt = 1.f - 2.f * rng_next(r0, r1);
            const float ss = 1.f - 2.f * rng_next(r0, r1);
            const float pp = sqrtf(fmaxf(1.f - vx * vx - vy * vy, 0.f)) * (rng_next(r0, r1) > 0.5f ? 1.f : -1.f);
            const float qx = vy * pp - vz * ss, qy = vz * tt - vx * pp, qz = vx * ss - vy * tt;
            vx = qx;
            vy = qy;
            vz = qz;
        }
        ox += P1.x * 0.5f;
        oy += P1.y * 0.5f;
        oz += P1.z * 0.5f;
        canfocus = (st == MMC_SRC_SLIT);
    }
    if (canfocus && D.w != 0.f) {
        /* converging (focus > 0) or diverging (focus < 0) beam */
        ox += D.w * vx;
        oy += D.w * vy;
        oz += D.w * vz;
        if (D.w < 0.f) {
            vx = px - ox;
            vy = py - oy;
            vz = pz - oz;
        } else {
            vx = ox - px;
            vy = oy - py;
            vz = oz - pz;
        }
        const float rn = 1.f / sqrtf(vx * vx + vy * vy + vz * vz);
        vx = vx * rn;
        vy = vy * rn;
        vz = vz * rn;
    }
    return true;
}

/* SSE4.1 dpps over xyz: (ax*bx + ay*by) + (az*bz + 0) */
__device__ __forceinline__ float dpps(float ax, float ay, float az, float bx, float by, float bz) {
    float t2 = ax * bx + ay * by;
    float t3 = az * bz + 0.0f;
    return t2 + t3;
}

/* Kahan add into a lane-private LDS pair */
__device__ __forceinline__ void kahan_lds(float* s, float* c, float v) {
    float y = v - *c;
    float t = *s + y;
    *c = (t - *s) - y;
    *s = t;
}

/* FP32 field deposit into this workgroup's copy of the field (f, lo). */
template <bool COMP>
__device__ __forceinline__ void deposit(float* f, float* lo, uint64_t idx, float v) {
    if (!COMP) {
        atomicAdd(f + idx, v);
    } else {
        /* TwoSum against the value the atomic returned: hi + lo holds the
         * exact running sum up to the rounding of lo itself */
        float old = atomicAdd(f + idx, v);
        float s = old + v;
        float bp = s - old;
        float err = (old - (s - bp)) + (v - bp);
        if (err != 0.f) atomicAdd(lo + idx, err);
    }
}

/* Weight before the step as the deposit sees it: the photon's own weight
 * before attenuation, or in replay the detected weight rw times the step's
 * Jacobian factor exp(-dmua L), path length L or scattering fraction
 * L mus / slen0, added to the attenuated weight w
 * (src/mmc_raytrace.c:1523-1543). */
__device__ __forceinline__ float replay_weight(int rmode, float currweight, float w, float L, float mus,
                                               float slen0, float rw) {
    if (rmode == kReplayNone) return currweight;
    float c;
    if (rmode == kReplayJ) c = mmc_expf(-DELTA_MUA * L);
    else if (rmode == kReplayWL) c = L;
    else c = (slen0 < EPSF) ? 1.f : L * mus / slen0;
    c *= rw;
    return c + w;
}

/* Physical XCD (accelerator complex die) executing this wave, 0..7 on MI355X. */
__device__ __forceinline__ uint32_t xcd_id() {
    uint32_t x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return x & 0xF;
}

/*
 * The transport kernel.  TR selects the Plucker tracer with linear nodal
 * deposits (src/mmc_raytrace.c:227-508), the branch-less Badouel tracer with
 * element deposits (src/mmc_raytrace.c:1414-1709), or the same tracer with
 * dual-grid voxel deposits (src/mmc_raytrace.c:1615-1673).
 */
template <int TR, bool COMP, int OCC, int TAB>
__global__ void __launch_bounds__(kBlock, OCC) transport(KArgs karg) {
    (void)karg;                                   /* read through kargs() */
    const KArgs* ka = kargs();
#define a (*ka)
    constexpr bool NODAL = TR == kTrNodal;        /* Plucker tracer, nodal deposits */
    /* branch-less Badouel tracer with the linear (nodal) basis: a third of
     * each step's deposit to each node of the face the ray leaves through */
    constexpr bool BNODE = TR == kTrElemNodal;
    constexpr bool GRID = TR == kTrGrid || TR == kTrGridRF;
    /* RF forward (omega > 0, mmc_core.cl IS_RF): complex photon weight and
     * complex fluence, imaginary part in a second block of the field */
    constexpr bool RF = TR == kTrElemRF || TR == kTrGridRF;
    /* per-simulation tables: in LDS (small batches) or through the caches
     * (large batches).  (Reading a single simulation's tables as wave-uniform
     * values was measured 16 % slower on the slab: the scalar register file
     * overflows into lanes of vector registers.) */
    constexpr bool STAGE = TAB == kTabLds;
    extern __shared__ __align__(16) unsigned char smem[];
    const int nsim = a.nsim;
    const int nmed = a.prop + 1 + a.isextdet;
    /* per-simulation tables: staged in LDS for batches whose tables are
     * small, read through L1/L2 otherwise (lanes of a wave mostly share a
     * simulation, so those reads broadcast) */
    const int nstage = STAGE ? nsim : 0;
    SimHot* s_hot = (SimHot*)smem;
    float4* s_med = (float4*)(s_hot + nstage);
    int64_t* s_cum = (int64_t*)(s_med + (size_t)nstage * nmed);
    float* s_lane = (float*)(s_cum + nsim + 1);
    const int nthr = blockDim.x;
    const int tid = threadIdx.x;
    /* lane-private LDS: [slot][thread] */
    auto LS = [&](int slot) -> float* { return s_lane + (size_t)slot * nthr + tid; };
    const int ppbase = kLaneSlots;                   /* nscat[prop], ppath[prop], mom[prop] */
    const int slot_w0 = kLaneSlots + (2 + a.mom) * a.prop;
    const int slot_rg = slot_w0 + 1;
    const int slot_ns = slot_w0 + 2;
    /* shared patterns (srcnum > 1): the photon's weight in each pattern, and
     * this lane's launched and absorbed energy per pattern */
    const int slot_pw = slot_w0 + 3;
    const int slot_pl = slot_pw + a.maxsrc;
    const int slot_pa = slot_pl + a.maxsrc;
    /* with wide-field detector elements: the record's last column as the
     * reference builds it (launch weight, or pixel bits, plus what lands
     * there from detector elements) */
    const int slot_wc = slot_w0 + 3 + (a.maxsrc > 1 ? 3 * a.maxsrc : 0);
    const int nslots = slot_wc + a.isextdet;

    for (int i = tid; i < nstage; i += nthr) s_hot[i] = a.hot[i];
    for (int i = tid; i < nstage * nmed; i += nthr) s_med[i] = a.media[i];
    for (int i = tid; i <= nsim; i += nthr) s_cum[i] = a.cum[i];
    for (int k = 0; k < nslots; k++) *LS(k) = 0.f;
    __shared__ float s_thr[kThrLds];
    /* the current element's geometry, per lane: Badouel's four face planes
     * (nx, ny, nz, d as float4 rows) or Plucker's 12 node coordinates; kept
     * while the photon scatters inside the element, reloaded after a
     * crossing */

    for (int i = tid; i < kThrLds; i += nthr) s_thr[i] = g_lm_thr[kThrLog + i];   /* log, azimuth */
    __syncthreads();

    const int lane = __lane_id();
    /* Each XCD has its own L2: with one field copy per XCD no cache line is
     * shared between dies, so the atomics stay local to one L2 instead of
     * contending across the fabric (all copies are summed in FP64 at the end). */
    /* this workgroup's field copy: its die's, and with more than 8 copies
     * (a multiple of 8) one of the die's nrep/8, taken by workgroup number */
    const uint32_t nrepu = (uint32_t)a.nrep;
    const uint64_t rep = nrepu <= 1 ? 0
                         : nrepu <= 8 ? (uint64_t)(xcd_id() % nrepu)
                                      : (uint64_t)(xcd_id() % 8u + 8u * ((blockIdx.x >> 3) % (nrepu >> 3)));
    float* const fcopy = a.field + rep * a.repstride;
    float* const lcopy = COMP ? a.field_lo + rep * a.repstride : nullptr;
    const uint64_t lanemask_lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    /* ---- lane state ---- */
    bool alive = false;
    bool exhausted = false;
    int64_t next_g = 0, end_g = 0;      /* this lane's claimed photon ids */
    int sim = -1;
    int64_t gid = 0;
    float px = 0.f, py = 0.f, pz = 0.f, vx = 0.f, vy = 0.f, vz = 1.f;
    int eid = 0, ctype = 0;
    float w = 0.f, tof = 0.f, slen = 0.f, Lmove = 0.f;
    float b0 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
    uint32_t oldidx = 0xFFFFFFFFu;
    float oldw = 0.f;
    uint64_t r0 = 0, r1 = 0;
    int ctx = CTX_OUTER, fixcount = 0;
    float eabs = 0.f, eabs_c = 0.f;
    uint32_t raytet = 0, raytet_start = 0;
    int cache_e = -1;                   /* element whose geometry is in cNX..cDD and cfnb */
    float4 cNX = make_float4(0.f, 0.f, 0.f, 0.f), cNY = cNX, cNZ = cNX, cDD = cNX;
    int4 cfnb = make_int4(0, 0, 0, 0), cee = cfnb;
    /* pending nodal deposits (Plucker) to the nodes of element cache_e at
     * field offset nbase */
    float na0 = 0.f, na1 = 0.f, na2 = 0.f, na3 = 0.f;
    uint64_t nfoff = 0;                 /* field offset of the simulation */
    uint32_t nloc = 0;                  /* gate * nn */
    bool npend = false;
    int nsrc = 1;                       /* patterns sharing the photon (field slots) */
    float rw = 1.f, rt = 0.f;           /* replay: detected weight and time of the photon */
    float wi = 0.f, oldwi = 0.f;        /* RF: imaginary weight and its coalescing register */
    float wm = 0.f;                     /* RF: weight magnitude (absorption ledger, roulette) */
    float slen0 = 0.f;                  /* scattering length at the last scattering event */
    int posidx = 0;                     /* launch pixel of a pattern source */
    /* The first two draws of the next scattering event (scattering length,
     * azimuth) are evaluated one event ahead, from a copy of the generator
     * state: their floats (pf_l, pf_s, pf_c), the directions of their far
     * neighbours and low bits of their generator outputs (pf_bits), and the
     * agreement table bytes (pf_tl, pf_ta) -- fetched only for draws near a
     * rounding midpoint, so that the fetch latency hides behind the steps in
     * between.  Valid while no other draw intervenes (reflection and roulette
     * draws invalidate them). */
    float pf_l = 0.f, pf_s = 0.f, pf_c = 1.f;
    uint32_t pf_tl = 0u, pf_ta = 0u, pf_bits = 0u;
    bool pf_ok = false;

    /* the lane's simulation index into the tables, and its hot parameters */
    auto sidx = [&]() -> int { return sim; };
    auto hot_of = [&]() -> SimHot { return STAGE ? s_hot[sim] : a.hot[sim]; };
    /* partial path of a step in medium ctype (src/mmc_raytrace.c:1869,
     * 1954, 1969): record column prop-1+ctype.  For a wide-field detector
     * element (ctype = prop+1) that is column 2 prop, which the reference
     * leaves where it lands: the first momentum column, else the exit
     * columns (rewritten when the photon leaves), else the launch-weight
     * column. */
    auto add_path = [&](float L) {
        if (ctype <= a.prop) {
            *LS(ppbase + a.prop - 1 + ctype) += L;
        } else if (hot_of().flags & kFlagMom) {
            *LS(ppbase + 2 * a.prop) += L;
        } else if (!a.cold[sim].issaveexit) {
            *LS(slot_wc) += L;
        }
    };

    /* deposit v at value loc of the simulation's field (offset foff); with
     * shared patterns the value has nsrc slots and slot k receives v times
     * the photon's weight in pattern k (src/mmc_raytrace.c:1586-1612) */
    auto depo = [&](uint64_t foff, uint32_t loc, float v) {
        if (nsrc == 1) {
            deposit<COMP>(fcopy, lcopy, foff + loc, v);
        } else {
            for (int k = 0; k < nsrc; k++) {
                const float pk = *LS(slot_pw + k);
                if (pk != 0.f) deposit<COMP>(fcopy, lcopy, foff + (uint64_t)loc * nsrc + k, v * pk);
            }
        }
    };
    auto flush_nodal = [&]() {
        /* x + 0 == x for the non-negative field: empty node sums are skipped */
        if (na0 != 0.f) depo(nfoff, nloc + (uint32_t)(cee.x - 1), na0);
        if (na1 != 0.f) depo(nfoff, nloc + (uint32_t)(cee.y - 1), na1);
        if (na2 != 0.f) depo(nfoff, nloc + (uint32_t)(cee.z - 1), na2);
        if (na3 != 0.f) depo(nfoff, nloc + (uint32_t)(cee.w - 1), na3);
        npend = false;
    };
    /* the geometry and neighbours of element e into the cache registers
     * (Badouel: the 4 planes; Plucker: the 12 node coordinates in cNX..cNZ
     * and the node ids) */
    auto load_elem = [&](int e) {
        if (!NODAL) {
            const float4* P = a.planes + 4 * (size_t)e;
            cNX = P[0]; cNY = P[1]; cNZ = P[2]; cDD = P[3];
            if (BNODE) cee = a.elem[e];
        } else {
            const float4* Q = a.enode + 3 * (size_t)e;
            cNX = Q[0]; cNY = Q[1]; cNZ = Q[2];
            cee = a.elem[e];
        }
        cfnb = a.facenb[e];
        cache_e = e;
    };
    /* one trajectory record: photon id, position, weight, element
     * (savedebugdata, src/mmc_raytrace.c:1739-1756) */
    auto savetraj = [&]() {
        const unsigned long long k = atomicAdd(a.detcount + nsim + 1, 1ull);
        if (k < a.maxtraj) {
            const SimCold& C = a.cold[sidx()];
            const uint32_t pid = (uint32_t)(C.photon_first + (gid - s_cum[sidx()]));
            float* r = a.traj + (size_t)k * MMC_TRAJ_RECLEN;
            r[0] = __uint_as_float(pid);
            r[1] = px;
            r[2] = py;
            r[3] = pz;
            r[4] = w;
            r[5] = __int_as_float(eid);
        }
    };
    auto flush_sim = [&](int s) {
        if (s < 0) return;
        double* acc = a.simacc + (size_t)s * kAccStride;
        float v;
        v = *LS(0) - *LS(1); if (v != 0.f) atomicAdd(acc + kAccLaunch, (double)v);
        v = *LS(2) - *LS(3); if (v != 0.f) atomicAdd(acc + kAccAbsorb, (double)v);
        v = *LS(4) - *LS(5); if (v != 0.f) atomicAdd(acc + kAccEscape, (double)v);
        v = *LS(6);          if (v != 0.f) atomicAdd(acc + kAccRGain, (double)v);
        if (raytet) atomicAdd(acc + kAccRaytet, (double)raytet);
        for (int k = 0; k < kLaneSlots; k++) *LS(k) = 0.f;
        if (a.maxsrc > 1)
            for (int k = 0; k < a.maxsrc; k++) {
                v = *LS(slot_pl + k); if (v != 0.f) atomicAdd(acc + kAccSrcLaunch + k, (double)v);
                v = *LS(slot_pa + k); if (v != 0.f) atomicAdd(acc + kAccSrcAbsorb + k, (double)v);
                *LS(slot_pl + k) = 0.f;
                *LS(slot_pa + k) = 0.f;
            }
        raytet = 0;
    };

    while (true) {
        ka = kargs();
        /* ------------------------------------------------ refill dead lanes */
        /* Lanes whose photon ended take the next id of their private range;
         * lanes with an empty range are refilled together: one atomic per
         * wave claims chunk ids for each of them. */
        const bool need = !alive && next_g >= end_g;
        const uint64_t needmask = __ballot(need);
        if (needmask && !exhausted) {
            const int nneed = __popcll(needmask);
            const int leader = __ffsll((unsigned long long)needmask) - 1;
            const unsigned long long want = (unsigned long long)nneed * (unsigned long long)a.chunk;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(a.queue, want);
            uint32_t blo = __shfl((uint32_t)base, leader);
            uint32_t bhi = __shfl((uint32_t)(base >> 32), leader);
            base = ((unsigned long long)bhi << 32) | blo;
            if ((int64_t)(base + want) >= a.total) exhausted = true;
            if (need) {
                next_g = (int64_t)base + (int64_t)__popcll(needmask & lanemask_lt) * a.chunk;
                end_g = min(next_g + (int64_t)a.chunk, a.total);
            }
        }
        {
            if (!alive && next_g < end_g) {
                {
                    const int64_t g = next_g++;
                    /* locate the simulation of photon g */
                    int lo = 0, hi = nsim - 1;
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (s_cum[mid] <= g) lo = mid; else hi = mid - 1;
                    }
                    if (lo != sim) {
                        flush_sim(sim);
                        sim = lo;
                    }
                    gid = g;
                    alive = true;
                    const SimCold& C = a.cold[sidx()];
                    uint64_t pid = (uint64_t)(C.photon_first + (g - s_cum[sidx()]));
                    if (C.replaybase >= 0) {
                        /* replay: the saved initial state of this photon */
                        const int64_t k = C.replaybase + (g - s_cum[sidx()]);
                        r0 = a.seeds[2 * k];
                        r1 = a.seeds[2 * k + 1];
                        rw = a.rweight ? a.rweight[k] : 1.f;
                        rt = a.rtime ? a.rtime[k] : 0.f;
                    } else {
                        photon_seed(C.seed, pid, r0, r1);
                        rw = 1.f;
                        rt = 0.f;
                    }
                    slen0 = 0.f;
                    nsrc = C.srcnum;
                    posidx = 0;
                    /* launchphoton (src/mmc_raytrace.c:2346-2671) */
                    px = C.srcpos.x; py = C.srcpos.y; pz = C.srcpos.z;
                    vx = C.srcdir.x; vy = C.srcdir.y; vz = C.srcdir.z;
                    b0 = C.bary0.x; b1 = C.bary0.y; b2 = C.bary0.z; b3 = C.bary0.w;
                    eid = C.e0;
                    w = 1.f; tof = 0.f; Lmove = 0.f;
                    oldidx = 0xFFFFFFFFu; oldw = 0.f;
                    if (RF) {
                        wi = 0.f;
                        oldwi = 0.f;
                    }
                    ctx = CTX_OUTER; fixcount = 0;
                    eabs = 0.f; eabs_c = 0.f;
                    raytet_start = raytet;
                    /* drop the element cache: a new photon reloads its first
                     * element once, and the launch code below (source
                     * geometry, angle evaluations in double) gets the cache's
                     * registers */
                    cache_e = -1;
                    cfnb = make_int4(0, 0, 0, 0);
                    cee = cfnb;
                    cNX = make_float4(0.f, 0.f, 0.f, 0.f);
                    cNY = cNX;
                    cNZ = cNX;
                    cDD = cNX;
                    na0 = na1 = na2 = na3 = 0.f;
                    slen = rng_scatlen(r0, r1, s_thr);
                    const bool search = launch_source(C, a.pattern, r0, r1, px, py, pz, vx, vy, vz, w, posidx, s_thr);
                    if (C.srctype == MMC_SRC_PATTERN && nsrc > 1)
                        for (int k = 0; k < nsrc; k++) {
                            const float pk = a.pattern[C.patbase + posidx * nsrc + k];
                            *LS(slot_pw + k) = pk;
                            *LS(slot_pl + k) += pk;
                        }
                    if (search) {
                        px = EPSF * vx + 1.f * px;
                        py = EPSF * vy + 1.f * py;
                        pz = EPSF * vz + 1.f * pz;
                        /* enclosing element: e0 first, then the candidate list in
                         * its order; with a locator grid only the candidates whose
                         * (widened) bounding box covers the launch point's cell are
                         * tested, in the same order, so the first match is the same */
                        /* candidates in order: k = -1 is e0, then either the
                         * locator cell's list or the whole srcelem list; one
                         * loop keeps a single copy of the containment test */
                        int kbeg = -1, kend = a.srcelemlen;
                        const bool useloc = a.loc_start != nullptr;
                        if (useloc) {
                            int cx = (int)floorf((px - a.loc_min.x) * a.loc_inv.x);
                            int cy = (int)floorf((py - a.loc_min.y) * a.loc_inv.y);
                            int cz = (int)floorf((pz - a.loc_min.z) * a.loc_inv.z);
                            /* clamped like the host's cell ranges (a monotone map) */
                            cx = min(max(cx, 0), a.loc_dim.x - 1);
                            cy = min(max(cy, 0), a.loc_dim.y - 1);
                            cz = min(max(cz, 0), a.loc_dim.z - 1);
                            int cell = (cz * a.loc_dim.y + cy) * a.loc_dim.x + cx;
                            kbeg = a.loc_start[cell] - 1;
                            kend = a.loc_start[cell + 1];
                        }
                        const int e0 = eid;
                        bool found = false;
                        for (int k = kbeg; k < kend && !found; k++) {
                            int ce;
                            if (k == kbeg) {
                                if (e0 <= 0 || e0 > a.ne) continue;
                                ce = e0;
                            } else {
                                ce = a.srcelem[useloc ? a.loc_elem[k] : k];
                            }
                            const int4 ee = a.elem[ce - 1];
                            const float4 N0 = a.node[ee.x - 1], N1 = a.node[ee.y - 1], N2 = a.node[ee.z - 1],
                                         N3 = a.node[ee.w - 1];
                            /* signed sub-volume opposite each node (src/mmc_raytrace.c:2566-2590) */
                            float bc[4];
                            #pragma unroll
                            for (int f = 0; f < 4; f++) {
                                const float4 A = f == 1 ? N3 : N0;
                                const float4 B = f == 0 ? N3 : (f == 1 ? N2 : (f == 2 ? N2 : N1));
                                const float4 Cc = f == 0 ? N1 : (f == 1 ? N1 : (f == 2 ? N3 : N2));
                                float abx = B.x - A.x, aby = B.y - A.y, abz = B.z - A.z;
                                float acx = Cc.x - A.x, acy = Cc.y - A.y, acz = Cc.z - A.z;
                                float sx = px - A.x, sy = py - A.y, sz = pz - A.z;
                                float nx = aby * acz - abz * acy;
                                float ny = abz * acx - abx * acz;
                                float nz = abx * acy - aby * acx;
                                bc[(0x2013 >> (4 * (3 - f))) & 0xF] = -(sx * nx + sy * ny + sz * nz);
                            }
                            if (bc[0] < -1e-4f || bc[1] < -1e-4f || bc[2] < -1e-4f || bc[3] < -1e-4f) continue;
                            float sum = 0.f;
                            sum += bc[0]; sum += bc[1]; sum += bc[2]; sum += bc[3];
                            eid = ce;
                            b0 = bc[0] / sum; b1 = bc[1] / sum; b2 = bc[2] / sum; b3 = bc[3] / sum;
                            found = true;
                        }
                        if (!found) eid = ID_UNDEF;   /* reported as a lost photon */
                    }
                    ctype = (eid > 0 && eid != ID_UNDEF) ? a.type[eid - 1] : 0;
                    *LS(slot_w0) = w;
                    if (RF) wm = w;
                    *LS(slot_rg) = 0.f;
                    *LS(slot_ns) = 0.f;
                    if (a.savedet)
                        for (int k = 0; k < (2 + a.mom) * a.prop; k++) *LS(ppbase + k) = 0.f;
                    if (a.isextdet) *LS(slot_wc) = (nsrc > 1) ? __int_as_float(posidx) : w;
                    /* replay with weighted path / scattering outputs counts the
                     * detected weight as launched energy (src/mmc_raytrace.c:1815);
                     * a pattern source first scales it by the launch pixel's
                     * pattern value (src/mmc_raytrace.c:2394-2399) */
                    {
                        const int rmode = (hot_of().flags >> kReplayShift) & 3;
                        if ((rmode == kReplayWL || rmode == kReplayWP) && C.srctype == MMC_SRC_PATTERN)
                            rw *= a.pattern[C.patbase + posidx];
                        kahan_lds(LS(0), LS(1), (rmode == kReplayWL || rmode == kReplayWP) ? rw : w);
                    }
                    if (eid == ID_UNDEF) ctx = -1;   /* ends below without tracing */
                    if (hot_of().flags & kFlagTraj) savetraj();
                    pf_ok = false;
                }
            }
        }
        if (__ballot(alive) == 0) break;
        if (!alive) continue;

        const SimHot hot = hot_of();
        const int rmode = (hot.flags >> kReplayShift) & 3;
        const size_t mbase = (size_t)sidx() * nmed;
        auto med = [&](int k) -> float4 { return STAGE ? s_med[mbase + k] : a.media[mbase + k]; };
        int reason = 0;
        int faceid = -1, isend = 0, undef = 1, nexteid = 0, faceidx = 4;
        float poutx = 0.f, pouty = 0.f, poutz = 0.f;

        if (ctx >= 0 && raytet - raytet_start >= a.maxsteps) {
            ctx = -1;                      /* bounded lifetime: ends as a lost photon */
            reason = MMC_END_STEPLIMIT;
        }
        if (ctx >= 0) {
            /* ============================ ray-tetrahedron step ============ */
            const int e = eid - 1;
            /* the element's geometry and neighbours stay in registers while
             * the photon scatters inside it; a crossing loads the next
             * element's, so only a photon's first element is loaded here */
            if (e != cache_e) load_elem(e);
            const int4 fnb = cfnb;
            const float4 m = med(ctype);
            const float mua = m.x, mus = m.y, nidx = m.w;
            const float rc = nidx * R_C0F;
            const float currweight = w;
            raytet++;
            if (!NODAL) {
                const float4 NX = cNX, NY = cNY, NZ = cNZ, DD = cDD;
                float T[4];
                #pragma unroll
                for (int i = 0; i < 4; i++) {
                    float nx = pick4f(NX, i), ny = pick4f(NY, i), nz = pick4f(NZ, i), d = pick4f(DD, i);
                    float t = nx * px;
                    t = t + ny * py;
                    t = t + nz * pz;
                    t = d - t;
                    float s = nx * vx;
                    s = s + ny * vy;
                    s = s + nz * vz;
                    t = t / s;
                    T[i] = (s > 0.f) ? t : 1e10f;
                }
                float m02 = (T[0] < T[2]) ? T[0] : T[2];
                float m13 = (T[1] < T[3]) ? T[1] : T[3];
                float bx = (m02 < m13) ? m02 : m13;
                faceidx = (T[3] == bx) ? 3 : (T[2] == bx) ? 2 : (T[1] == bx) ? 1 : (T[0] == bx) ? 0 : 4;
                faceid = (faceidx < 4) ? faceorder(faceidx) : -1;
                if (faceid >= 0 && bx >= 0.f) {
                    nexteid = pick4(fnb, faceid);
                    float dlen = (mus <= EPSF) ? R_MIN_MUSF : slen / mus;
                    isend = (bx > dlen);
                    Lmove = isend ? dlen : bx;
                    poutx = vx * bx + px;
                    pouty = vy * bx + py;
                    poutz = vz * bx + pz;
                    undef = 0;
                    if ((int)((tof + Lmove * rc - hot.tstart) * hot.rtstep) > hot.maxgate - 1) {
                        faceid = -2;
                        undef = 1;
                        Lmove = (hot.tend - tof) / (nidx * R_C0F) - 1e-4f;
                    }
                    const float att = mmc_expf(-mua * Lmove);
                    w *= att;
                    slen -= Lmove * mus;
                    const float cw = replay_weight(rmode, currweight, w, Lmove, mus, slen0, rw);
                    float ww = cw - w;
                    tof += Lmove * rc;
                    int g = (rmode == kReplayWL || rmode == kReplayWP) ? (int)(rt * hot.rtstep)
                                                                       : (int)((tof - hot.tstart) * hot.rtstep);
                    const int framelen = GRID ? a.grid[sidx()].ncells : (BNODE ? a.nn : a.ne);
                    int tshift = max(min(g, hot.maxgate - 1), 0) * framelen;
                    float dabs = ww;
                    if (RF) {
                        /* the magnitude of the complex weight decays as the
                         * real weight does without modulation */
                        const float wb = wm;
                        wm *= att;
                        dabs = wb - wm;
                    }
                    if (mua > 0.f) {
                        float y = dabs - eabs_c, tt = eabs + y;
                        eabs_c = (tt - eabs) - y;
                        eabs = tt;
                    }
                    if (!(hot.flags & kFlagEnergy))
                        ww = (mua < EPSF) ? (cw * Lmove) : ((cw - w) / mua);
                    /* RF: complex Beer-Lambert over the step, w <- w e^{-(mua + i k) L}
                     * with k = omega n / c0, and the complex deposit
                     * (w_before - w_after) / (mua + i k) (src/mmc_core.cl:868-897) */
                    float ww_im = 0.f;
                    const float kap = RF ? a.omega * nidx * R_C0F : 0.f;
                    const float amag2 = mua * mua + kap * kap;
                    if (RF && !GRID) {
                        const float tl = 1.f - att;
                        const float attr = (tl < 1.f) ? (1.f - tl) : 1.f;
                        const float phase = a.omega * nidx * R_C0F * Lmove;
                        float sphi, cphi;
                        mmc_sincosf(phase, &sphi, &cphi);
                        const float w0r = currweight, w0i = wi;
                        const float nre = attr * (w0r * cphi + w0i * sphi);
                        const float nim = attr * (-w0r * sphi + w0i * cphi);
                        const float dwr = w0r - nre, dwi = w0i - nim;
                        ww = (amag2 > 0.f) ? (dwr * mua + dwi * kap) / amag2 : (w0r * Lmove);
                        ww_im = (amag2 > 0.f) ? (dwi * mua - dwr * kap) / amag2 : (w0i * Lmove);
                        w = nre;
                        wi = nim;
                    }
                    const uint64_t imoff = RF ? (uint64_t)framelen * (uint64_t)hot.maxgate : 0;
                    const float q0x = px, q0y = py, q0z = pz;
                    px = px + vx * Lmove;
                    py = py + vy * Lmove;
                    pz = pz + vz * Lmove;
                    const uint64_t foff = hot.fieldoff;
                    if (BNODE) {
                        /* linear basis under the Badouel tracer
                         * (src/mmc_raytrace.c:1674-1692): ww/3 to each node
                         * of the exit face out[faceidx] = {0,3,1} {3,2,1}
                         * {0,2,3} {0,1,2}, i.e. every node but the one
                         * opposite (2, 0, 1, 3); the thirds of consecutive
                         * steps in one element and gate add up in registers
                         * and reach the field when the photon leaves the
                         * element, changes gate or ends */
                        const uint32_t loc = (uint32_t)tshift;
                        const float w3 = ww * (1.f / 3.f);
                        const int opp = nib4(0x2013, faceidx);
                        if (npend && loc != nloc) flush_nodal();
                        if (!npend) {
                            na0 = na1 = na2 = na3 = 0.f;
                            nfoff = foff;
                            nloc = loc;
                            npend = true;
                        }
                        if (opp != 0) na0 += w3;
                        if (opp != 1) na1 += w3;
                        if (opp != 2) na2 += w3;
                        if (opp != 3) na3 += w3;
                        /* (at a crossing the sums of the exit face's nodes,
                         * which the next element shares, follow the photon;
                         * see the crossing below) */
                    } else if (!GRID) {
                        uint32_t newidx = (uint32_t)(e + tshift);
                        if (oldidx == 0xFFFFFFFFu) oldidx = newidx;
                        if (newidx != oldidx) {
                            /* the first step after a face crossing finds the
                             * register already flushed (oldw == 0): the
                             * reference adds that zero (src/mmc_raytrace.c:
                             * 1562-1612); x + 0 == x for the non-negative
                             * field, so the atomic is skipped, which halves
                             * the atomics.  RF: the GPU reference flushes the
                             * register only when its real part is positive
                             * (src/mmc_core.cl:905) */
                            if (RF) {
                                if (oldw > 0.f) {
                                    depo(foff, oldidx, oldw);
                                    if (oldwi != 0.f) depo(foff + imoff, oldidx, oldwi);
                                }
                                oldwi = ww_im;
                            } else if (oldw != 0.f) {
                                depo(foff, oldidx, oldw);
                            }
                            oldidx = newidx;
                            oldw = ww;
                        } else {
                            oldw += ww;
                            if (RF) oldwi += ww_im;
                        }
                        if (faceid == -2 || !isend) {
                            if (RF) {
                                if (oldw != 0.f) depo(foff, newidx, oldw);
                                if (oldwi != 0.f) depo(foff + imoff, newidx, oldwi);
                                oldwi = 0.f;
                            } else {
                                depo(foff, newidx, oldw);
                            }
                            oldw = 0.f;
                        }
                    } else {
                        /* dual grid (src/mmc_raytrace.c:1615-1673): the step
                         * is cut into an even number of segments no longer
                         * than half a voxel; segment i deposits the fraction
                         * segloss^i (1 - segloss) / (1 - att) of the step's
                         * deposit at the voxel holding its midpoint (clamped
                         * to the grid), coalesced in the same register */
                        const SimGrid G = a.grid[sidx()];
                        const float tl = 1.f - att;
                        int seg = (int)(Lmove / G.step) + 1;
                        seg = seg << 1;
                        const float dstep = Lmove / (float)seg;
                        const float segloss = mmc_expf(-mua * dstep);
                        const float tx = vx * dstep, ty = vy * dstep, tz = vz * dstep;
                        float sx = (q0x - G.org.x) + tx * 0.5f;
                        float sy = (q0y - G.org.y) + ty * 0.5f;
                        float sz = (q0z - G.org.z) + tz * 0.5f;
                        const float frac = (tl == 0.f) ? 0.f : (1.f - segloss) / tl;
                        const bool flush = (faceid == -2 || !isend);
                        float w0 = ww;
                        /* RF: the complex weight turns by the same phase in
                         * every segment (src/mmc_core.cl:1043-1081) */
                        float sr = currweight, si = wi, scos = 1.f, ssin = 0.f;
                        if (RF) mmc_sincosf(a.omega * nidx * R_C0F * dstep, &ssin, &scos);
                        for (int i = 0; i < seg; i++) {
                            const int ix = min(max((int)(sx * G.org.w), 0), G.nx - 1);
                            const int iy = min(max((int)(sy * G.org.w), 0), G.ny - 1);
                            const int iz = min(max((int)(sz * G.org.w), 0), G.nz - 1);
                            const uint32_t newidx = (uint32_t)(iz * G.dxy + iy * G.dx + ix + tshift);
                            if (oldidx == 0xFFFFFFFFu) oldidx = newidx;
                            float d = w0 * frac, dim = 0.f;
                            if (RF) {
                                const float w0r = sr, w0i = si;
                                const float nre = segloss * (w0r * scos + w0i * ssin);
                                const float nim = segloss * (-w0r * ssin + w0i * scos);
                                sr = nre;
                                si = nim;
                                const float dwr = w0r - nre, dwi = w0i - nim;
                                d = (amag2 < EPSF) ? (w0r * w0) : (dwr * mua + dwi * kap) / amag2;
                                dim = (amag2 < EPSF) ? (w0i * w0) : (dwi * mua - dwr * kap) / amag2;
                            }
                            if (newidx != oldidx) {
                                if (oldw != 0.f) depo(foff, oldidx, oldw);
                                if (RF) {
                                    if (oldwi != 0.f) depo(foff + imoff, oldidx, oldwi);
                                    oldwi = dim;
                                }
                                oldidx = newidx;
                                oldw = d;
                            } else {
                                oldw += d;
                                if (RF) oldwi += dim;
                            }
                            if (flush) {
                                if (oldw != 0.f) depo(foff, newidx, oldw);
                                oldw = 0.f;
                                if (RF) {
                                    if (oldwi != 0.f) depo(foff + imoff, newidx, oldwi);
                                    oldwi = 0.f;
                                }
                            }
                            w0 *= segloss;
                            sx = sx + tx;
                            sy = sy + ty;
                            sz = sz + tz;
                        }
                        if (RF) {
                            w = sr;
                            wi = si;
                        }
                    }
                }
                if (faceid == -2) slen = 0.f;
            } else {
                /* Plucker test over the 6 edges (src/mmc_raytrace.c:227-508) */
                const int4 ee = cee;
                const float4 q0 = cNX, q1 = cNY, q2 = cNZ;
                const float nxs[4] = {q0.x, q0.w, q1.z, q2.y};
                const float nys[4] = {q0.y, q1.x, q1.w, q2.z};
                const float nzs[4] = {q0.z, q1.y, q2.x, q2.w};
                float p1x = px + vx, p1y = py + vy, p1z = pz + vz;
                float cx = py * p1z - pz * p1y;
                float cy = pz * p1x - px * p1z;
                float cz = px * p1y - py * p1x;
                float wv[6];
                #pragma unroll
                for (int i = 0; i < 6; i++) {
                    const int i0 = (i < 3) ? 0 : (i < 5 ? 1 : 2);
                    const int i1 = (i == 0) ? 1 : (i == 1 || i == 3) ? 2 : 3;
                    float ax = nxs[i0], ay = nys[i0], az = nzs[i0];
                    float bxx = nxs[i1], byy = nys[i1], bzz = nzs[i1];
                    float dx = bxx - ax, dy = byy - ay, dz = bzz - az;
                    float mx = ay * bzz - az * byy, my = az * bxx - ax * bzz, mz = ax * byy - ay * bxx;
                    wv[i] = dpps(vx, vy, vz, mx, my, mz) + dpps(cx, cy, cz, dx, dy, dz);
                }
                float bo[4] = {0.f, 0.f, 0.f, 0.f};
                float Lp0 = 0.f;
                Lmove = 0.f;
                /* the four faces in order; face i sees Plucker products
                 * (a, b, c) = wv[fc[i]] with b negated for faces 2 and 3, and
                 * the ray leaves through the first face with a < 0, b < 0,
                 * c >= 0 by sign bit.  All four are tested and the exit
                 * computed once for the selected face. */
                const float fa[4] = {wv[0], wv[3], wv[2], wv[1]};
                const float fb[4] = {wv[4], wv[5], -wv[5], -wv[3]};
                const float fcv[4] = {wv[2], wv[4], wv[1], wv[0]};
                int sel = 4;
                #pragma unroll
                for (int i = 3; i >= 0; i--) {
                    uint32_t ua = __float_as_uint(fa[i]), ub = __float_as_uint(fb[i]), uc = __float_as_uint(fcv[i]);
                    if ((ua & ub & ~uc) >> 31) sel = i;
                }
                if (sel < 4) {
                    const float wa = sel == 0 ? fa[0] : (sel == 1 ? fa[1] : (sel == 2 ? fa[2] : fa[3]));
                    const float wb = sel == 0 ? fb[0] : (sel == 1 ? fb[1] : (sel == 2 ? fb[2] : fb[3]));
                    const float wc = sel == 0 ? fcv[0] : (sel == 1 ? fcv[1] : (sel == 2 ? fcv[2] : fcv[3]));
                    /* face nodes nc[i] = {3,0,1} {3,1,2} {2,0,3} {1,0,2} */
                    const int na = nib4(0x3321, sel), nb = nib4(0x0100, sel), nc = nib4(0x1232, sel);
                    float Rv = 1.f / (-wa - wb + wc);
                    float ba = -wa * Rv, bb = -wb * Rv, bcc = wc * Rv;
                    #pragma unroll
                    for (int k = 0; k < 4; k++) bo[k] = (k == na) ? ba : ((k == nb) ? bb : ((k == nc) ? bcc : 0.f));
                    const float xa = pick4a(nxs, na), xb = pick4a(nxs, nb), xc = pick4a(nxs, nc);
                    const float ya = pick4a(nys, na), yb = pick4a(nys, nb), yc = pick4a(nys, nc);
                    const float za = pick4a(nzs, na), zb = pick4a(nzs, nb), zc = pick4a(nzs, nc);
                    poutx = ba * xa + bb * xb + bcc * xc;
                    pouty = ba * ya + bb * yb + bcc * yc;
                    poutz = ba * za + bb * zb + bcc * zc;
                    undef = 0;
                    float dx = poutx - px, dy = pouty - py, dz = poutz - pz;
                    Lp0 = sqrtf(dx * dx + dy * dy + dz * dz);
                    float dlen = (mus <= EPSF) ? R_MIN_MUSF : slen / mus;
                    faceidx = sel;
                    faceid = faceorder(sel);
                    nexteid = pick4(fnb, faceid);
                    isend = (Lp0 > dlen);
                    Lmove = isend ? dlen : Lp0;
                }
                if (!undef) {
                    if ((int)((tof + Lmove * rc - hot.tstart) * hot.rtstep) >=
                        (int)((hot.tend - hot.tstart) * hot.rtstep)) {
                        faceid = -2;
                        undef = 1;
                        Lmove = (hot.tend - tof) / (nidx * R_C0F) - 1e-4f;
                    }
                    w *= mmc_expf(-mua * Lmove);
                    slen -= Lmove * mus;
                    px += Lmove * vx;
                    py += Lmove * vy;
                    pz += Lmove * vz;
                    if (Lp0 > EPSF) {
                        tof += Lmove * rc;
                        const float cw = replay_weight(rmode, currweight, w, Lmove, mus, slen0, rw);
                        float ww = cw - w;
                        if (mua > 0.f || !(hot.flags & kFlagEnergy)) {
                            float ratio = Lmove / Lp0;
                            if (mua > 0.f) {
                                float y = ww - eabs_c, tt = eabs + y;
                                eabs_c = (tt - eabs) - y;
                                eabs = tt;
                            }
                            if (!(hot.flags & kFlagEnergy))
                                ww = (mua < EPSF) ? (cw * Lmove) : ((cw - w) / mua);
                            int g = (rmode == kReplayWL || rmode == kReplayWP)
                                        ? (int)(rt * hot.rtstep)
                                        : (int)((tof - hot.tstart) * hot.rtstep);
                            const uint32_t loc = (uint32_t)max(min(g, hot.maxgate - 1), 0) * (uint32_t)a.nn;
                            ww *= 0.5f;
                            if (isend) {
                                bo[0] = (1.f - ratio) * b0 + ratio * bo[0];
                                bo[1] = (1.f - ratio) * b1 + ratio * bo[1];
                                bo[2] = (1.f - ratio) * b2 + ratio * bo[2];
                                bo[3] = (1.f - ratio) * b3 + ratio * bo[3];
                            }
                            /* the four nodal deposits of consecutive steps in
                             * one element and gate add up in registers and
                             * reach the field when the photon leaves the
                             * element, changes gate or ends */
                            const float d0 = ww * (b0 + bo[0]), d1 = ww * (b1 + bo[1]);
                            const float d2 = ww * (b2 + bo[2]), d3 = ww * (b3 + bo[3]);
                            if (npend && loc != nloc) flush_nodal();
                            if (npend) {
                                na0 += d0; na1 += d1; na2 += d2; na3 += d3;
                            } else {
                                na0 = d0; na1 = d1; na2 = d2; na3 = d3;
                                nfoff = hot.fieldoff;
                                nloc = loc;
                                npend = true;
                            }
                        }
                        /* (the sums of the nodes shared with the next
                         * element follow the photon across the face; see
                         * the crossing below) */
                        if (isend) {
                            b0 = bo[0]; b1 = bo[1]; b2 = bo[2]; b3 = bo[3];
                        } else if (nb_id(nexteid, a.lblshift) > 0 && faceidx < 4) {
                            /* carry the exit-face barycentrics into the next element */
                            int4 ne4 = a.elem[nb_id(nexteid, a.lblshift) - 1];
                            float nbv[4] = {0.f, 0.f, 0.f, 0.f};
                            #pragma unroll
                            for (int j = 0; j < 3; j++) {
                                const int nj = nib4(j == 0 ? 0x3321 : (j == 1 ? 0x0100 : 0x1232), faceidx);
                                int node = pick4(ee, nj);
                                float bv = pick4a(bo, nj);
                                if (ne4.x == node) nbv[0] = bv;
                                else if (ne4.y == node) nbv[1] = bv;
                                else if (ne4.z == node) nbv[2] = bv;
                                else if (ne4.w == node) nbv[3] = bv;
                            }
                            b0 = nbv[0]; b1 = nbv[1]; b2 = nbv[2]; b3 = nbv[3];
                        }
                    }
                    if (faceid == -2) slen = 0.f;
                }
            }

            /* ======================= onephoton bookkeeping after the test == */
            const bool savedet = (hot.flags & kFlagSaveDet) != 0;
            bool next_iter = false;
            if (ctx == CTX_OUTER) {
                if (undef) {
                    if (fixcount++ < MAX_TRIAL) {
                        ctx = CTX_OUTER;
                        next_iter = true;
                    } else {
                        eid = ID_UNDEF;
                        faceid = -1;
                    }
                }
                if (!next_iter && savedet && Lmove > 0.f && faceid >= 0 && ctype > 0) add_path(Lmove);
            } else {
                if (savedet && Lmove > 0.f && ctype > 0) add_path(Lmove);
                if (ctx == CTX_INNER) fixcount = 0;
                if (undef) {
                    if (fixcount++ < MAX_TRIAL) {
                        ctx = CTX_INNERFIX;
                        next_iter = true;
                    } else {
                        eid = ID_UNDEF;
                    }
                }
            }
            if (next_iter) {
                /* fixphoton: nudge 0.1% toward the centroid (src/mmc_raytrace.c:200-213) */
                int4 ee = a.elem[eid - 1];
                float4 n0 = a.node[ee.x - 1], n1 = a.node[ee.y - 1], n2 = a.node[ee.z - 1], n3 = a.node[ee.w - 1];
                float cxx = 0.f, cyy = 0.f, czz = 0.f;
                cxx = cxx + n0.x; cyy = cyy + n0.y; czz = czz + n0.z;
                cxx = cxx + n1.x; cyy = cyy + n1.y; czz = czz + n1.z;
                cxx = cxx + n2.x; cyy = cyy + n2.y; czz = czz + n2.z;
                cxx = cxx + n3.x; cyy = cyy + n3.y; czz = czz + n3.z;
                px += (cxx * 0.25f - px) * FIX_PHOTON;
                py += (cyy * 0.25f - py) * FIX_PHOTON;
                pz += (czz * 0.25f - pz) * FIX_PHOTON;
                continue;
            }

            /* ---- crossing into the neighbour (inner while of onephoton) ---- */
            if (eid != ID_UNDEF && faceid >= 0 && !isend) {
                const int isreflect = (hot.flags >> kReflShift) & 3;
                const int isspecular = (hot.flags >> kSpecShift) & 3;
                px = poutx; py = pouty; pz = poutz;
                int oldeid = eid;
                int oldtype = ctype;
                int nbp = nexteid;
                int neid = nb_id(nbp, a.lblshift);
                int ntype = neid > 0 ? (a.lblshift ? (nbp & 15) : a.type[neid - 1]) : 0;
                float nold = med(oldtype).w;
                if (isreflect && (neid <= 0 || med(ntype).w != nold)) {
                    if (!(neid <= 0 && ((nold == hot.nout && isreflect != MMC_BC_MIRROR) ||
                                        isreflect == MMC_BC_ABSORBEXT))) {
                        /* reflectray (src/mmc_raytrace.c:2257-2332) */
                        int f = ifaceorder(faceid);
                        const float4* P = a.planes + 4 * (size_t)(oldeid - 1);
                        float pnx = pick4f(P[0], f), pny = pick4f(P[1], f), pnz = pick4f(P[2], f);
                        float Icos = fabsf(vx * pnx + vy * pny + vz * pnz);
                        float n1 = nold;
                        float n2 = (neid > 0) ? med(ntype).w : hot.nout;
                        float t0 = n1 * n1, t1 = n2 * n2;
                        float t2 = 1.f - t0 / t1 * (1.f - Icos * Icos);
                        if (t2 > 0.f && !(neid <= 0 && isreflect == MMC_BC_MIRROR)) {
                            float Re = t0 * Icos * Icos + t1 * t2;
                            t2 = sqrtf(t2);
                            float Im = 2.f * n1 * n2 * Icos * t2;
                            float Rt = (Re - Im) / (Re + Im);
                            Re = t1 * Icos * Icos + t0 * t2 * t2;
                            Rt = (Rt + (Re - Im) / (Re + Im)) * 0.5f;
                            pf_ok = false;
                            if (rng_next(r0, r1) <= Rt) {
                                float k = -2.f * Icos;
                                vx = 1.f * vx + k * pnx; vy = 1.f * vy + k * pny; vz = 1.f * vz + k * pnz;
                                neid = oldeid;
                                ntype = oldtype;
                            } else if (isspecular == 2 && neid == 0) {
                            } else {
                                float k = -Icos;
                                vx = 1.f * vx + k * pnx; vy = 1.f * vy + k * pny; vz = 1.f * vz + k * pnz;
                                float r = n1 / n2;
                                vx = r * vx + t2 * pnx; vy = r * vy + t2 * pny; vz = r * vz + t2 * pnz;
                            }
                        } else {
                            float k = -2.f * Icos;
                            vx = 1.f * vx + k * pnx; vy = 1.f * vy + k * pny; vz = 1.f * vz + k * pnz;
                            neid = oldeid;
                            ntype = oldtype;
                        }
                        float inv = 1.f / sqrtf(vx * vx + vy * vy + vz * vz);
                        vx = inv * vx; vy = inv * vy; vz = inv * vz;
                    }
                }
                eid = neid;
                if (eid > 0) {
                    if (oldtype == 0 && ntype != 0 && !(hot.flags & kFlagVoidTime)) tof = 0.f;
                    /* leaving the tissue into the background ends the photon
                     * unless wide-field detectors wait outside */
                    if (oldtype != 0 && ntype == 0 && !a.isextdet) eid = 0;
                }
                if (eid > 0) {
                  if (eid != oldeid) {
                    /* the next element's geometry is fetched here, so the
                     * loads overlap the bookkeeping up to its first step */
                    const int4 oe = cee;
                    load_elem(eid - 1);
                    if ((NODAL || BNODE) && npend) {
                        /* Nodal sums across the face: the three nodes
                         * the next element shares keep their sums (in the next
                         * element's node order, which cee takes when its
                         * geometry is loaded); the node left behind is
                         * deposited.  Same sums per node as a deposit at every
                         * crossing, a quarter of the atomics. */
                        const int4 nn = cee;
                        float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
                        const int on[4] = {oe.x, oe.y, oe.z, oe.w};
                        const float ov[4] = {na0, na1, na2, na3};
                        #pragma unroll
                        for (int i = 0; i < 4; i++) {
                            if (nn.x == on[i]) m0 = ov[i];
                            else if (nn.y == on[i]) m1 = ov[i];
                            else if (nn.z == on[i]) m2 = ov[i];
                            else if (nn.w == on[i]) m3 = ov[i];
                            else if (ov[i] != 0.f) depo(nfoff, nloc + (uint32_t)(on[i] - 1), ov[i]);
                        }
                        na0 = m0; na1 = m1; na2 = m2; na3 = m3;
                    }
                  }
                    ctype = ntype;
                    ctx = CTX_INNER;
                    continue;
                }
            }

            /* ---- exit, roulette, scattering ---- */
            if (eid <= 0 || eid == ID_UNDEF || undef) {
                reason = (eid == ID_UNDEF) ? MMC_END_LOST : MMC_END_EXIT;
            } else {
                /* RF: the roulette tests the magnitude of the complex weight
                 * and scales both parts (src/mmc_core.cl:2090-2110) */
                if ((hot.flags & kFlagRoulette) && (RF ? wm : w) < hot.minenergy) {
                    const float rsize = hot.roulettesize;
                    pf_ok = false;
                    if (rng_next(r0, r1) * rsize <= 1.f) {
                        float wb = RF ? wm : w;
                        w *= rsize;
                        if (RF) {
                            wi *= rsize;
                            wm *= rsize;
                        }
                        *LS(slot_rg) += (RF ? wm : w) - wb;
                    } else {
                        reason = MMC_END_ROULETTE;
                    }
                }
                if (!reason) {
                    /* mc_next_scatter (src/mmc_mesh.c:1644-1715) */
                    const float g = med(ctype).z;
                    float stheta = 0.f, ctheta = 1.f;
                    /* pass 0 evaluates this event's first two draws when
                     * they were not evaluated ahead; pass 1 applies them,
                     * draws the zenith, turns the direction and evaluates
                     * the next event's first two draws */
                    #pragma unroll 1
                    for (int pass = pf_ok ? 1 : 0; pass < 2; pass++) {
                        if (pass == 1) {
                            (void)rng_next(r0, r1);          /* the two draws evaluated ahead */
                            (void)rng_next(r0, r1);
                            const float uz = rng_next(r0, r1);
                            const uint32_t kl = pf_bits & 7u, ka = (pf_bits >> 3) & 3u;
                            const float nextslen =
                                -(((pf_tl >> kl) & 1u) ? mmc_fstep(pf_l, (pf_bits >> 5) & 1u) : pf_l);
                            const float sphi = ((pf_ta >> (2 * ka)) & 1u) ? mmc_fstep(pf_s, (pf_bits >> 6) & 1u) : pf_s;
                            const float cphi =
                                ((pf_ta >> (2 * ka + 1)) & 1u) ? mmc_fstep(pf_c, (pf_bits >> 7) & 1u) : pf_c;
                            float tmp0, tmp1;
                            if (g > EPSF) {
                                tmp0 = (1.f - g * g) / (1.f - g + 2.f * g * uz);
                                tmp0 *= tmp0;
                                tmp0 = (1.f + g * g - tmp0) / (2.f * g);
                                tmp0 = (tmp0 > 1.f) ? 1.f : tmp0;
                                tmp0 = (tmp0 < -1.f) ? -1.f : tmp0;
                                stheta = sqrtf(1.f - tmp0 * tmp0);
                                ctheta = tmp0;
                            } else {
                                angle_draw(kAngIso, uz, &stheta, &ctheta, s_thr);
                            }
                            float qx, qy, qz;
                            if (vz > -1.f + EPSF && vz < 1.f - EPSF) {
                                tmp0 = 1.f - vz * vz;
                                tmp1 = 1.f / sqrtf(tmp0);
                                tmp1 = stheta * tmp1;
                                qx = tmp1 * (vx * vz * cphi - vy * sphi) + vx * ctheta;
                                qy = tmp1 * (vy * vz * cphi + vx * sphi) + vy * ctheta;
                                qz = -tmp1 * tmp0 * cphi + vz * ctheta;
                            } else {
                                qx = stheta * cphi;
                                qy = stheta * sphi;
                                qz = (vz > 0.f) ? ctheta : -ctheta;
                            }
                            vx = qx; vy = qy; vz = qz;
                            slen = nextslen;
                            slen0 = nextslen;
                        }
                        /* the next two draws from a copy of the state:
                         * -logf(u + EPS) and sincosf((float)(2 pi u)) in
                         * double, rounded, with the table bytes of those near
                         * a midpoint (src/mmc_rand_common.h:50-51,
                         * src/mmc_mesh.c:1662-1666) */
                        uint64_t a0 = r0, a1 = r1;
                        const float u1 = rng_next(a0, a1), u2 = rng_next(a0, a1);
                        const uint32_t k1 = draw_bits(u1), k2 = draw_bits(u2);
                        const double el = mmc_log_dk(u1 + EPSF, dk());
                        double es, ec;
                        mmc_sincos_dk(mmc_twopi(u2), &es, &ec, dk());
                        pf_l = (float)el;
                        pf_s = (float)es;
                        pf_c = (float)ec;
                        const float* tha = s_thr + MMC_LIBM_NREG + 2 * (k2 >> MMC_LIBM_REGSHIFT);
                        const bool nl = lm_near(el, pf_l, s_thr[k1 >> MMC_LIBM_REGSHIFT]);
                        const bool na = lm_near(es, pf_s, tha[0]) || lm_near(ec, pf_c, tha[1]);
                        pf_tl = nl ? (uint32_t)g_lm_log[k1 >> 3] : 0u;
                        pf_ta = na ? (uint32_t)g_lm_az[k2 >> 2] : 0u;
                        pf_bits = (k1 & 7u) | ((k2 & 3u) << 3) | ((uint32_t)(el > (double)pf_l) << 5) |
                                  ((uint32_t)(es > (double)pf_s) << 6) | ((uint32_t)(ec > (double)pf_c) << 7);
                        pf_ok = true;
                    }
                    if (hot.flags & kFlagTraj) savetraj();
                    if (savedet && ctype > 0) *LS(ppbase + ctype - 1) += 1.f;
                    /* momentum transfer (src/mmc_raytrace.c:2073-2075) */
                    if ((hot.flags & kFlagMom) && ctype > 0) {
                        if (ctype <= a.prop) *LS(ppbase + 2 * a.prop - 1 + ctype) += 1.f - ctheta;
                        else if (!a.cold[sidx()].issaveexit) *LS(slot_wc) += 1.f - ctheta;
                    }
                    *LS(slot_ns) += 1.f;
                    ctx = CTX_OUTER;
                    continue;
                }
            }
        } else if (!reason) {
            reason = MMC_END_LOST;
        }

        /* ============================================ photon ends here ===== */
        {
            if ((NODAL || BNODE) && npend) flush_nodal();
            if (hot.flags & kFlagTraj) savetraj();
            const SimCold& C = a.cold[sidx()];
            const float w0 = *LS(slot_w0);
            int exitdet = 0;
            if (reason == MMC_END_EXIT && a.savedet && C.issavedet && eid <= 0 && a.isextdet &&
                C.detnum == 0 && ctype == a.prop + 1) {
                /* left the mesh through a wide-field detector element: its
                 * id is the detector id (src/mmc_raytrace.c:2013-2014) */
                exitdet = cache_e + 1;
            }
            if (reason == MMC_END_EXIT && a.savedet && C.issavedet && eid <= 0) {
                for (int i = 0; i < C.detnum && !exitdet; i++) {
                    float4 d = a.detpos[C.detbase + i];
                    if ((d.x - px) * (d.x - px) + (d.y - py) * (d.y - py) + (d.z - pz) * (d.z - pz) <
                        d.w * d.w) {
                        exitdet = i + 1;
                        break;
                    }
                }
                if (exitdet) {
                    unsigned long long slot = atomicAdd(a.detcount, 1ull);
                    atomicAdd(a.detcount + 1 + sim, 1ull);
                    if ((int64_t)slot < a.maxdet && a.detdata) {
                        float* row = a.detdata + (size_t)slot * a.reclen;
                        row[0] = (float)exitdet;
                        const int npp = (hot_of().flags & kFlagMom) ? 3 * a.prop : 2 * a.prop;
                        for (int k = 0; k < npp; k++) row[1 + k] = *LS(ppbase + k);
                        int col = 1 + npp;
                        if (C.issaveexit) {
                            row[col++] = px; row[col++] = py; row[col++] = pz;
                            row[col++] = vx; row[col++] = vy; row[col++] = vz;
                        }
                        /* shared patterns record the launch pixel instead of
                         * the launch weight (src/mmc_raytrace.c:1824) */
                        row[col] = a.isextdet ? *LS(slot_wc) : (nsrc > 1) ? __int_as_float(posidx) : w0;
                        if (a.detphoton) a.detphoton[slot] = gid;
                    }
                }
            }
            /* RF: the escaping energy is the weight's magnitude
             * (src/mmc_core.cl:2145-2153) */
            const float wend = (reason == MMC_END_ROULETTE) ? 0.f : (RF ? wm : w);
            float rg = *LS(slot_rg);
            if (reason == MMC_END_ROULETTE) rg -= (RF ? wm : w);
            kahan_lds(LS(2), LS(3), eabs);
            if (nsrc > 1)
                for (int k = 0; k < nsrc; k++) *LS(slot_pa + k) += eabs * *LS(slot_pw + k);
            kahan_lds(LS(4), LS(5), wend);
            *LS(6) += rg;
            if (a.rec && gid >= a.recfirst && gid < a.recfirst + a.reccount) {
                mmc_hip_photon* R = a.rec + (gid - a.recfirst);
                R->p[0] = px; R->p[1] = py; R->p[2] = pz; R->w = wend;
                R->v[0] = vx; R->v[1] = vy; R->v[2] = vz; R->t = tof;
                R->eid = eid; R->reason = reason;
                R->nraytet = (int32_t)(raytet - raytet_start);
                R->detid = exitdet;
                R->eabs = eabs; R->rgain = rg; R->w0 = w0; R->nscat = *LS(slot_ns);
            }
            alive = false;
        }
    }
    flush_sim(sim);
}

/* FP32 -> FP64 accumulation of a finished field (the final reduction): the
 * nrep copies (and their compensation buffers) summed in copy order; zeroes
 * the FP32 buffers for reuse.  out may be null when only the sum is wanted. */
#undef a

__global__ void field_reduce(float* f, float* lo, double* out, uint64_t n, double* sum, int nrep,
                             uint64_t repstride) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    double part = 0.0;
    for (; i < n; i += stride) {
        double v = 0.0;
        for (int r = 0; r < nrep; r++) {
            const uint64_t k = i + (uint64_t)r * repstride;
            v += (double)f[k];
            f[k] = 0.f;
            if (lo) {
                v += (double)lo[k];
                lo[k] = 0.f;
            }
        }
        if (out) out[i] = v;
        part += v;
    }
    /* wave reduction, one FP64 atomic per wave */
    for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off);
    if ((threadIdx.x & 63) == 0 && part != 0.0) atomicAdd(sum, part);
}

__global__ void fill_zero_u64(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

#define MMC_INSTANTIATE(OCC, TAB)                                           \
    template __global__ void transport<kTrElem, false, OCC, TAB>(KArgs);    \
    template __global__ void transport<kTrElem, true, OCC, TAB>(KArgs);     \
    template __global__ void transport<kTrNodal, false, OCC, TAB>(KArgs);   \
    template __global__ void transport<kTrNodal, true, OCC, TAB>(KArgs);    \
    template __global__ void transport<kTrGrid, false, OCC, TAB>(KArgs);    \
    template __global__ void transport<kTrGrid, true, OCC, TAB>(KArgs);      \
    template __global__ void transport<kTrElemRF, false, OCC, TAB>(KArgs);  \
    template __global__ void transport<kTrElemRF, true, OCC, TAB>(KArgs);   \
    template __global__ void transport<kTrGridRF, false, OCC, TAB>(KArgs);  \
    template __global__ void transport<kTrGridRF, true, OCC, TAB>(KArgs);   \
    template __global__ void transport<kTrElemNodal, false, OCC, TAB>(KArgs); \
    template __global__ void transport<kTrElemNodal, true, OCC, TAB>(KArgs);
MMC_INSTANTIATE(3, kTabLds)
MMC_INSTANTIATE(4, kTabLds)
MMC_INSTANTIATE(5, kTabLds)
MMC_INSTANTIATE(4, kTabGlobal)

}  // namespace mmc
