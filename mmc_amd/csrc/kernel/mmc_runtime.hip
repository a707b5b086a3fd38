/*
 * mmc_runtime.hip -- host side of the C ABI in mmc_hip.h: device mesh
 * residency, simulation tables, the persistent launch, the FP64 final
 * reduction and the copies back to the caller.
 *
 * Plays the role of mmc_run_cl (src/mmc_cl_host.c:87-1721) / mmc_run_cu
 * (src/mmc_cu_host.cu:203-1554): buffer build, launch, read-back and
 * accumulation into the caller's double field.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <chrono>
#include <memory>
#include <thread>
#include <vector>

#include "mmc_hip.h"
#include "mmc_kernel.h"

using namespace mmc;

static thread_local char g_err[512] = "";

static int set_err(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

/* the message mmc_hip_last_error returns, for the other entry points */
int mmc::report_error(int code, const char* msg) { return set_err(code, "%s", msg); }

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            return set_err(-100 - (int)e_, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                           \
    } while (0)

struct mmc_hip_meshres {
    int nn, ne, prop, nf;
    int isextdet = 0;             /* elements labelled prop+1 are wide-field detectors */
    int lblshift = 4;             /* facenb packing: 4 (labels in the low bits) or 0 */
    float4* planes = nullptr;
    int4* facenb = nullptr;
    int4* elem = nullptr;
    int32_t* type = nullptr;
    float4* node = nullptr;
    float4* enode = nullptr;
    int32_t* srcelem = nullptr;
    int32_t srcelemlen = 0;
    int32_t* loc_start = nullptr;
    int32_t* loc_elem = nullptr;
    float4 loc_min, loc_inv;
    int4 loc_dim;
    std::vector<float> med;       /* (prop+1)*4 default media */
    float gmin[3], gmax[3];       /* node bounding box widened by EPS (dual grid) */
    int64_t bytes = 0;
};

extern "C" int32_t mmc_hip_abi_version(void) { return MMC_HIP_ABI_VERSION; }
extern "C" const char* mmc_hip_last_error(void) { return g_err; }

extern "C" int32_t mmc_hip_device_info(char* name, int32_t len, int64_t* hbm_bytes) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_err(-1, "no HIP device");
        return -1;
    }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return -1;
    if (name && len > 0) snprintf(name, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
    if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
    return p.multiProcessorCount;
}

extern "C" void mmc_hip_photon_seeds(uint64_t seed, int64_t first, int64_t n, uint64_t* out) {
    for (int64_t i = 0; i < n; i++) photon_seed(seed, (uint64_t)(first + i), out[2 * i], out[2 * i + 1]);
}

/* ---- host C library agreement tables --------------------------------------
 * Built once per process from the C library this process runs with
 * (mmc_host_libm_table) and copied once into each device's copy of the
 * kernel's tables, before that device's first launch.  Host copies are
 * static, so later launches allocate nothing for them. */
static std::mutex g_libm_mu;
static uint8_t g_tab_bits[4][MMC_FIXB_ISO];
static float g_tab_thr[4][MMC_LIBM_NREG * 3];
static bool g_tab_built = false;
static bool g_tab_dev[64];

static int ensure_libm_tables() {
    std::lock_guard<std::mutex> lk(g_libm_mu);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_err(-1, "device index %d out of range", dev);
    if (g_tab_dev[dev]) return 0;
    if (!g_tab_built) {
        if (!__builtin_cpu_supports("fma"))
            return set_err(-1, "the host CPU lacks FMA, needed to build the C library agreement tables");
        for (int w = 0; w < 4; w++) {
            const int e = mmc_host_libm_table(w, g_tab_bits[w], g_tab_thr[w]);
            if (e != 0)
                return set_err(-1, "C library agreement table %d: %d values the tables cannot reproduce", w, e);
        }
        g_tab_built = true;
    }
    int rc = upload_libm_tables(g_tab_bits[MMC_FIX_LOG], g_tab_bits[MMC_FIX_AZ], g_tab_bits[MMC_FIX_ISO],
                                g_tab_bits[MMC_FIX_PI], g_tab_thr[MMC_FIX_LOG], g_tab_thr[MMC_FIX_AZ],
                                g_tab_thr[MMC_FIX_ISO], g_tab_thr[MMC_FIX_PI]);
    if (rc) return rc;
    g_tab_dev[dev] = true;
    return 0;
}

/* dst[k] += src[k] over a field copied back from the device: split over up
 * to 8 host threads when the field is large (a memory-bound loop that one
 * thread runs at a fraction of the host's bandwidth); serial when threads
 * cannot be started. */
static void add_into(double* dst, const double* src, uint64_t n) {
    const uint64_t kMinPerThread = 1ull << 22;
    uint64_t nt = std::min<uint64_t>(std::max(1u, std::thread::hardware_concurrency()), 8);
    nt = std::min<uint64_t>(nt, (n + kMinPerThread - 1) / kMinPerThread);
    auto part = [dst, src, n, nt](uint64_t t) {
        const uint64_t b = n * t / nt, e = n * (t + 1) / nt;
        for (uint64_t k = b; k < e; k++) dst[k] += src[k];
    };
    if (nt > 1) {
        std::vector<std::thread> th;
        uint64_t started = 0;
        try {
            for (uint64_t t = 1; t < nt; t++, started++) th.emplace_back(part, t);
        } catch (...) {
        }
        part(0);
        for (auto& x : th) x.join();
        for (uint64_t t = started + 1; t < nt; t++) part(t);   /* parts whose thread did not start */
        return;
    }
    for (uint64_t k = 0; k < n; k++) dst[k] += src[k];
}

template <typename T>
static int dalloc(T** p, size_t count, int64_t* bytes) {
    size_t b = sizeof(T) * (count ? count : 1);
    hipError_t e = hipMalloc((void**)p, b);
    if (e != hipSuccess)
        return set_err(-2, "hipMalloc of %zu bytes failed: %s", b, hipGetErrorString(e));
    if (bytes) *bytes += (int64_t)b;
    return 0;
}

static void mesh_free(mmc_hip_meshres* m) {
    if (!m) return;
    (void)hipFree(m->planes);
    (void)hipFree(m->facenb);
    (void)hipFree(m->elem);
    (void)hipFree(m->type);
    (void)hipFree(m->node);
    (void)hipFree(m->enode);
    (void)hipFree(m->srcelem);
    (void)hipFree(m->loc_start);
    (void)hipFree(m->loc_elem);
    delete m;
}

/* Uniform grid over the candidate launch elements: each cell lists, in
 * ascending candidate order, the candidates whose bounding box (widened by a
 * margin that covers the -1e-4 barycentric tolerance of the launch test)
 * overlaps it.  Cell coordinates are computed with the same float operations
 * as the kernel's, so a point inside a widened box always falls in a cell of
 * that box's range. */
static hipError_t build_locator(const mmc_hip_mesh* mesh, mmc_hip_meshres* m) {
    const int n = mesh->srcelemlen;
    std::vector<float> lo(3 * (size_t)n), hi(3 * (size_t)n);
    float gmin[3] = {1e30f, 1e30f, 1e30f}, gmax[3] = {-1e30f, -1e30f, -1e30f};
    double mean_ext[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        const int32_t* e = mesh->elem + 4 * (size_t)(mesh->srcelem[i] - 1);
        for (int k = 0; k < 3; k++) {
            float a = 1e30f, b = -1e30f;
            for (int j = 0; j < 4; j++) {
                float v = mesh->node[3 * (size_t)(e[j] - 1) + k];
                a = std::min(a, v);
                b = std::max(b, v);
            }
            float pad = 1e-3f * (b - a) + 1e-3f;
            lo[3 * i + k] = a - pad;
            hi[3 * i + k] = b + pad;
            gmin[k] = std::min(gmin[k], a - pad);
            gmax[k] = std::max(gmax[k], b + pad);
            mean_ext[k] += (double)(b - a) / n;
        }
    }
    int dim[3];
    float inv[3];
    for (int k = 0; k < 3; k++) {
        double ext = std::max((double)gmax[k] - gmin[k], 1e-6);
        double cell = std::max(mean_ext[k], ext / 256.0);
        dim[k] = std::max(1, std::min(256, (int)std::ceil(ext / cell)));
        inv[k] = (float)(dim[k] / ext);
    }
    auto cell_of = [&](float v, int k) {
        int c = (int)floorf((v - gmin[k]) * inv[k]);
        return std::max(0, std::min(dim[k] - 1, c));
    };
    const size_t ncell = (size_t)dim[0] * dim[1] * dim[2];
    std::vector<int32_t> start(ncell + 1, 0);
    for (int pass = 0; pass < 2; pass++) {
        std::vector<int32_t> fill;
        if (pass) fill.assign(start.begin(), start.end() - 1);
        std::vector<int32_t>* out = nullptr;
        static thread_local std::vector<int32_t> items;
        if (pass) {
            items.assign((size_t)start[ncell], 0);
            out = &items;
        }
        for (int i = 0; i < n; i++) {
            int c0[3], c1[3];
            for (int k = 0; k < 3; k++) {
                c0[k] = cell_of(lo[3 * i + k], k);
                c1[k] = cell_of(hi[3 * i + k], k);
            }
            for (int z = c0[2]; z <= c1[2]; z++)
                for (int y = c0[1]; y <= c1[1]; y++)
                    for (int x = c0[0]; x <= c1[0]; x++) {
                        size_t c = ((size_t)z * dim[1] + y) * dim[0] + x;
                        if (!pass) start[c + 1]++;
                        else (*out)[fill[c]++] = i;
                    }
        }
        if (!pass) {
            for (size_t c = 0; c < ncell; c++) start[c + 1] += start[c];
        } else {
            int64_t dummy = 0;
            if (dalloc(&m->loc_start, ncell + 1, &m->bytes) || dalloc(&m->loc_elem, items.size(), &dummy))
                return hipErrorOutOfMemory;
            hipError_t e = hipMemcpy(m->loc_start, start.data(), sizeof(int32_t) * (ncell + 1), hipMemcpyHostToDevice);
            if (!e) e = hipMemcpy(m->loc_elem, items.data(), sizeof(int32_t) * items.size(), hipMemcpyHostToDevice);
            if (e) return e;
        }
    }
    m->loc_min = make_float4(gmin[0], gmin[1], gmin[2], 0.f);
    m->loc_inv = make_float4(inv[0], inv[1], inv[2], 0.f);
    m->loc_dim = make_int4(dim[0], dim[1], dim[2], 0);
    return hipSuccess;
}

extern "C" int32_t mmc_hip_mesh_upload(const mmc_hip_mesh* mesh, mmc_hip_mesh_handle* out) {
    if (!mesh || !out) return set_err(-1, "null argument");
    if (mesh->ne <= 0 || mesh->nn <= 0 || !mesh->node || !mesh->elem || !mesh->type ||
        !mesh->facenb || !mesh->planes || !mesh->med)
        return set_err(-1, "incomplete mesh");
    if (mesh->isextdet != 0 && mesh->isextdet != 1) return set_err(-1, "isextdet must be 0 or 1");
    if (mesh->prop < 1 || mesh->prop + mesh->isextdet > kMaxProp)
        return set_err(-1, "media count %d (+%d detector label) outside 1..%d", mesh->prop, mesh->isextdet, kMaxProp);
    /* the label of the next element travels in the 4 low bits of the packed
     * face neighbour while labels fit and ids leave room; otherwise the
     * neighbour ids are stored as they are and labels come from type */
    const bool packed = mesh->prop + mesh->isextdet <= kPackedProp && (int64_t)mesh->ne < (1ll << 27) &&
                        (int64_t)mesh->nf < (1ll << 27);
    const size_t ne = (size_t)mesh->ne, nn = (size_t)mesh->nn;
    /* validate ids on the host before anything reaches the device */
    for (size_t i = 0; i < 4 * ne; i++) {
        if (mesh->elem[i] < 1 || mesh->elem[i] > mesh->nn)
            return set_err(-1, "elem[%zu]=%d outside 1..nn", i, mesh->elem[i]);
        if (mesh->facenb[i] > mesh->ne || mesh->facenb[i] < -mesh->nf)
            return set_err(-1, "facenb[%zu]=%d outside -nf..ne", i, mesh->facenb[i]);
    }
    for (size_t i = 0; i < ne; i++)
        if (mesh->type[i] < 0 || mesh->type[i] > mesh->prop + mesh->isextdet)
            return set_err(-1, "type[%zu]=%d outside 0..prop%s", i, mesh->type[i], mesh->isextdet ? "+1" : "");

    mmc_hip_meshres* m = new mmc_hip_meshres();
    m->nn = mesh->nn;
    m->ne = mesh->ne;
    m->prop = mesh->prop;
    m->nf = mesh->nf;
    m->isextdet = mesh->isextdet;
    m->lblshift = packed ? 4 : 0;
    m->med.assign(mesh->med, mesh->med + 4 * (mesh->prop + 1));
    /* the dual grid's extent (mesh_createdualmesh, src/mmc_mesh.c:349-372):
     * the node bounding box widened by EPS; every grid starts at its corner */
    for (int k = 0; k < 3; k++) {
        float lo = 1e30f, hi = -1e30f;
        for (size_t i = 0; i < nn; i++) {
            const float v = mesh->node[3 * i + k];
            lo = (v < lo) ? v : lo;
            hi = (v > hi) ? v : hi;
        }
        m->gmin[k] = lo - 1e-6f;
        m->gmax[k] = hi + 1e-6f;
    }

    std::vector<int4> fnb(ne);
    std::vector<float4> nd(nn), en(3 * ne);
    for (size_t i = 0; i < ne; i++) {
        int v[4];
        for (int j = 0; j < 4; j++) {
            int nb = mesh->facenb[4 * i + j];
            int t = nb > 0 ? mesh->type[nb - 1] : 0;
            v[j] = packed ? ((int)((unsigned)nb << 4) | t) : nb;
        }
        fnb[i] = make_int4(v[0], v[1], v[2], v[3]);
    }
    for (size_t i = 0; i < nn; i++)
        nd[i] = make_float4(mesh->node[3 * i], mesh->node[3 * i + 1], mesh->node[3 * i + 2], 0.f);
    for (size_t i = 0; i < ne; i++) {
        float c[12];
        for (int j = 0; j < 4; j++) {
            const float* p = mesh->node + 3 * (size_t)(mesh->elem[4 * i + j] - 1);
            c[3 * j] = p[0];
            c[3 * j + 1] = p[1];
            c[3 * j + 2] = p[2];
        }
        en[3 * i] = make_float4(c[0], c[1], c[2], c[3]);
        en[3 * i + 1] = make_float4(c[4], c[5], c[6], c[7]);
        en[3 * i + 2] = make_float4(c[8], c[9], c[10], c[11]);
    }
    int rc = 0;
    if ((rc = dalloc(&m->planes, 4 * ne, &m->bytes)) || (rc = dalloc(&m->facenb, ne, &m->bytes)) ||
        (rc = dalloc(&m->elem, ne, &m->bytes)) || (rc = dalloc(&m->type, ne, &m->bytes)) ||
        (rc = dalloc(&m->node, nn, &m->bytes)) || (rc = dalloc(&m->enode, 3 * ne, &m->bytes)) ||
        (rc = dalloc(&m->srcelem, (size_t)std::max(mesh->srcelemlen, 1), &m->bytes))) {
        mesh_free(m);
        return rc;
    }
    hipError_t e = hipSuccess;
    e = e ? e : hipMemcpy(m->planes, mesh->planes, 64 * ne, hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(m->facenb, fnb.data(), 16 * ne, hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(m->elem, mesh->elem, 16 * ne, hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(m->type, mesh->type, 4 * ne, hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(m->node, nd.data(), 16 * nn, hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(m->enode, en.data(), 48 * ne, hipMemcpyHostToDevice);
    if (!e && mesh->srcelemlen > 0) {
        for (int i = 0; i < mesh->srcelemlen; i++)
            if (mesh->srcelem[i] < 1 || mesh->srcelem[i] > mesh->ne) {
                mesh_free(m);
                return set_err(-1, "srcelem[%d] outside 1..ne", i);
            }
        e = hipMemcpy(m->srcelem, mesh->srcelem, 4 * (size_t)mesh->srcelemlen, hipMemcpyHostToDevice);
    }
    m->srcelemlen = mesh->srcelemlen;
    if (!e && mesh->srcelemlen > 64) e = build_locator(mesh, m);
    if (e != hipSuccess) {
        mesh_free(m);
        return set_err(-3, "mesh upload failed: %s", hipGetErrorString(e));
    }
    *out = m;
    return 0;
}

extern "C" int32_t mmc_hip_mesh_release(mmc_hip_mesh_handle h) {
    mesh_free(h);
    return 0;
}

extern "C" int64_t mmc_hip_mesh_bytes(mmc_hip_mesh_handle h) { return h ? h->bytes : 0; }

/* field slots per value: the patterns that share each photon */
static int slots_of(const mmc_hip_sim* s) { return s->srctype == MMC_SRC_PATTERN ? s->srcnum : 1; }
/* momentum transfer is recorded with the detected photons only */
static int mom_of(const mmc_hip_sim* s, const mmc_hip_meshres* m) {
    return (s->ismomentum > 0 && s->issavedet && (s->detnum > 0 || m->isextdet)) ? 1 : 0;
}

/* dual grid of a simulation: mesh_createdualmesh (src/mmc_mesh.c:349-381)
 * over the mesh's widened box with the simulation's voxel edge */
static SimGrid grid_of(const mmc_hip_meshres* m, const mmc_hip_sim* s) {
    SimGrid g;
    memset(&g, 0, sizeof(g));
    const float step = s->gridstep;
    int dim[3];
    for (int k = 0; k < 3; k++) {
        float ext = m->gmax[k] - m->gmin[k];
        float q = ext / step;
        dim[k] = (q >= 0.f && q < 2.0e9f) ? (int)q + 1 : 0;
    }
    g.org = make_float4(m->gmin[0], m->gmin[1], m->gmin[2], 1.f / step);
    g.step = step;
    g.nx = dim[0];
    g.ny = dim[1];
    g.nz = dim[2];
    const int64_t dxy = (int64_t)dim[0] * dim[1], nc = dxy * dim[2];
    g.dx = dim[0];
    g.dxy = (dxy < (1ll << 31)) ? (int32_t)dxy : 0;
    g.ncells = (nc < (1ll << 31)) ? (int32_t)nc : 0;
    return g;
}

/* RF forward: complex weights, imaginary field block after the real one */
static bool rf_of(const mmc_hip_sim* s) { return s->omega > 0.f && !s->photonseed; }

/* values per gate (nodes, elements or voxels) times the pattern slots */
static int64_t datalen_of(const mmc_hip_meshres* m, const mmc_hip_sim* s) {
    if (s->method == MMC_RT_BLBADOUELGRID) return (int64_t)grid_of(m, s).ncells * slots_of(s);
    return (int64_t)(s->basisorder ? m->nn : m->ne) * slots_of(s);
}

extern "C" int32_t mmc_hip_sim_grid(mmc_hip_mesh_handle h, const mmc_hip_sim* s, float nmin[3], int32_t dim[3]) {
    if (!h || !s) return set_err(-1, "null argument");
    if (s->method != MMC_RT_BLBADOUELGRID || !(s->gridstep > 0.f))
        return set_err(-1, "not a dual-grid simulation (method %d, grid step %g)", s->method, (double)s->gridstep);
    SimGrid g = grid_of(h, s);
    if (!g.ncells) return set_err(-1, "grid step %g gives more than 2^31 voxels", (double)s->gridstep);
    if (nmin)
        for (int k = 0; k < 3; k++) nmin[k] = h->gmin[k];
    if (dim) {
        dim[0] = g.nx;
        dim[1] = g.ny;
        dim[2] = g.nz;
    }
    return 0;
}

extern "C" int64_t mmc_hip_sim_output_bytes(mmc_hip_mesh_handle h, const mmc_hip_sim* s) {
    if (!h || !s) return -1;
    /* FP32 working field + FP64 reduction target */
    return datalen_of(h, s) * (int64_t)s->maxgate * 12 * (rf_of(s) ? 2 : 1);
}

/* Host checks of a simulation against the mesh; every id the kernel will
 * dereference is validated here. */
static int check_sim(const mmc_hip_meshres* m, const mmc_hip_sim* s, int method0) {
    if (s->nphoton < 0) return set_err(-1, "negative photon count");
    if (s->method != MMC_RT_PLUCKER && s->method != MMC_RT_BLBADOUEL && s->method != MMC_RT_BLBADOUELGRID)
        return set_err(-1, "unsupported ray-tracer %d (use Plucker, branch-less Badouel or the dual grid)", s->method);
    if (s->method != method0) return set_err(-1, "all simulations of a batch must use one ray-tracer");
    if (s->method == MMC_RT_BLBADOUELGRID) {
        /* the dual grid always deposits per voxel (src/mmc_utils.c:3554-3556) */
        if (!(s->gridstep > 0.f)) return set_err(-1, "dual grid needs a positive voxel edge (gridstep)");
        const SimGrid g = grid_of(m, s);
        if (!g.ncells || (uint64_t)g.ncells * (uint64_t)s->maxgate >= (1ull << 31))
            return set_err(-1, "dual grid of %g-unit voxels too large (%d x %d x %d voxels, %d gates)",
                           (double)s->gridstep, g.nx, g.ny, g.nz, s->maxgate);
    } else if (s->method == MMC_RT_PLUCKER && s->basisorder != 1) {
        return set_err(-1, "the Plucker tracer deposits on nodes (basisorder 1)");
    } else if (s->basisorder != 0 && s->basisorder != 1) {
        return set_err(-1, "basisorder must be 0 (elements) or 1 (nodes)");
    } else {
        /* the kernel indexes a pattern's field (gate x element or node) in 32 bits */
        const uint64_t rows = s->basisorder == 1 ? (uint64_t)m->nn : (uint64_t)m->ne;
        if (s->maxgate > 0 && rows * (uint64_t)s->maxgate >= (1ull << 31))
            return set_err(-1, "field of %llu %s x %d gates exceeds 2^31 values", (unsigned long long)rows,
                           s->basisorder == 1 ? "nodes" : "elements", s->maxgate);
    }
    if (s->srctype < MMC_SRC_PENCIL || s->srctype > MMC_SRC_RING)
        return set_err(-1, "unsupported source type %d", s->srctype);
    if (s->srctype == MMC_SRC_PATTERN3D) {
        const float* p1 = s->srcparam1;
        if (!s->srcpattern || !(p1[0] >= 1.f && p1[1] >= 1.f && p1[2] >= 1.f) ||
            (double)p1[0] * p1[1] * p1[2] > 2e9)
            return set_err(-1, "3-D pattern source needs a pattern of srcparam1.x*y*z (>= 1 each) voxels");
    }
    if (s->srcnum > 1 && s->srctype != MMC_SRC_PATTERN)
        return set_err(-1, "shared patterns (srcnum %d) need a pattern source", s->srcnum);
    if (s->srctype == MMC_SRC_PENCILARRAY && !(s->srcparam1[3] >= 2.f && s->srcparam2[3] >= 2.f))
        return set_err(-1, "pencil-array source needs at least 2 beams along each axis (srcparam1.w, srcparam2.w)");
    if (s->srctype == MMC_SRC_CONE) {
        /* the half angle's rejection loop accepts with probability
         * (1 - cos a)/2, or a/(2 pi) for srcparam1.y > 0; keep it >= 1e-3 */
        const double ha = s->srcparam1[0];
        const double acc = s->srcparam1[1] > 0.f ? ha / 6.283185307179586 : (1.0 - cos(ha)) * 0.5;
        if (!(acc >= 1e-3))
            return set_err(-1, "cone source half angle %g too small (needs >= %s)", ha,
                           s->srcparam1[1] > 0.f ? "0.00629" : "0.0633 rad");
    }
    if (s->e0 < 1 || s->e0 > m->ne) return set_err(-1, "initial element %d outside 1..ne", s->e0);
    if (s->maxgate < 1 || !(s->tstep > 0.f) || !(s->tend > s->tstart))
        return set_err(-1, "bad time gates");
    if (s->srctype == MMC_SRC_PATTERN) {
        int nx = (int)s->srcparam1[3], ny = (int)s->srcparam2[3];
        if (!s->srcpattern || nx < 1 || ny < 1) return set_err(-1, "pattern source without pattern");
        if (s->srcnum < 1 || s->srcnum > kMaxSrc)
            return set_err(-1, "srcnum %d outside 1..%d patterns", s->srcnum, kMaxSrc);
    }
    if (s->issavedet && s->detnum > 0 && !s->detpos) return set_err(-1, "detectors without positions");
    if (s->outputtype < MMC_OUT_FLUX || s->outputtype > MMC_OUT_WP)
        return set_err(-1, "unsupported output type %d", s->outputtype);
    if (s->outputtype >= MMC_OUT_JACOBIAN) {
        if (!s->photonseed || !s->replayweight)
            return set_err(-1, "replay outputs need the photons' seeds and detected weights");
        if (s->outputtype != MMC_OUT_JACOBIAN && !s->replaytime)
            return set_err(-1, "weighted path/scattering outputs need the detected times");
        if (s->outputtype != MMC_OUT_JACOBIAN && slots_of(s) > 1)
            return set_err(-1, "replay of the weighted path/scattering outputs takes one source pattern "
                               "(src/mmc_raytrace.c:2396)");
    }
    if (!(s->omega >= 0.f)) return set_err(-1, "omega must be >= 0");
    if (rf_of(s) && (s->method == MMC_RT_PLUCKER || s->basisorder == 1 || slots_of(s) != 1))
        return set_err(-1, "RF (omega > 0) needs the branch-less Badouel tracer (element or grid output) and one "
                           "pattern");
    if (s->isreflect < MMC_BC_NOREFLECT || s->isreflect > MMC_BC_MIRROR)
        return set_err(-1, "unsupported boundary condition %d", s->isreflect);
    if (s->isspecular < 0 || s->isspecular > 2) return set_err(-1, "isspecular must be 0, 1 or 2");
    return 0;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

/* Device buffers that persist across launches and only grow; or, for the
 * allocation-free launch, a window of the static scratch array (fixed:
 * never allocated or freed here). */
struct Pooled {
    void* p = nullptr;
    size_t cap = 0;
    bool fixed = false;
    int ensure(size_t bytes) {
        if (bytes <= cap && p) return 0;
        if (fixed) return set_err(-1, "internal: scratch window of %zu bytes asked for %zu", cap, bytes);
        /* grow geometrically from 64 KiB: a free can wait for the device, so
         * steady-state launches must not reallocate */
        size_t want = std::max<size_t>(bytes, 64 << 10);
        if (cap) want = std::max(want, 2 * cap);
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        bytes = want;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return set_err(-2, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        cap = bytes;
        return 0;
    }
    ~Pooled() {
        if (p && !fixed) (void)hipFree(p);
    }
};

struct mmc_hip_arenares {
    float* field = nullptr;          /* nrep copies of n floats */
    float* lo = nullptr;
    uint64_t n = 0;
    int nrep = 1;
    /* collection: its own non-blocking stream and grow-only buffers, so a
     * collect neither waits for launches in flight nor frees device memory */
    hipStream_t st = nullptr;
    Pooled d64, dsum;
};

/* Field copies for a launch: one per XCD (8 on MI355X) while the copies fit
 * half of the free device memory and 64 GiB, else 4, 2 or 1. */
static int auto_replicas(uint64_t bytes_one_copy) {
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return 1;
    /* small fields take up to 8 copies per die: deposits near the source
     * then spread over more lines (the 60 mm cube's nodal field, 120 KB:
     * 1.38x with 64 copies over 8); the copies stay within 256 MiB */
    for (int r = kMaxReplicas; r > 8; r >>= 1)
        if ((uint64_t)r * bytes_one_copy <= (256ull << 20)) return r;
    const uint64_t budget = std::min<uint64_t>((uint64_t)freeb / 2, 64ull << 30);
    for (int r = 8; r > 1; r >>= 1)
        if ((uint64_t)r * bytes_one_copy <= budget) return r;
    return 1;
}

/* A launch slot: a stream, its timing events, the per-launch device buffers
 * (reused by the next launch on the slot) and the host state of the launch
 * in flight, which slot_finish needs to read the results back. */
struct mmc_hip_slotres {
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Pooled d_hot, d_cold, d_med, d_cum, d_det, d_pat, d_acc, d_detcnt, d_queue, d_detdata, d_detph, d_rec;
    Pooled d_seeds, d_rw, d_rt, d_grid, d_traj;
    /* launch in flight */
    bool busy = false;
    mmc_hip_mesh_handle m = nullptr;
    mmc_hip_arenares* arena = nullptr;
    int nsim = 0, nrep = 1, reclen = 0, savedet = 0;
    bool comp = false, has_rec = false;
    int64_t maxdet = 0, recfirst = 0, reccount = 0;
    int trajres = -1;                 /* result that receives the trajectory records */
    int64_t maxtraj = 0;
    uint64_t fieldlen = 0;
    std::vector<int64_t> cum;
    std::vector<uint64_t> fieldoff;
    std::vector<int64_t> datalen, maxdetph, photon_offset, nphoton;
    std::vector<int32_t> maxgate, nslot, rfmul;
    DevBuf field, lo;                 /* own fields of a launch without an arena */
    /* host sources of the launch's asynchronous copies, alive until it ends */
    std::vector<SimHot> h_hot;
    std::vector<SimCold> h_cold;
    std::vector<float4> h_media, h_dets;
    std::vector<float> h_pats;
    std::vector<uint64_t> h_seeds;
    std::vector<float> h_rw, h_rt;
    std::vector<SimGrid> h_grid;
    std::chrono::steady_clock::time_point t_start;
    /* allocation-free launches: the per-launch buffers are carved from this
     * device's scratch array (bump_base, bump_cap bytes) instead of pooled
     * allocations; the null stream, no events */
    unsigned char* bump_base = nullptr;
    size_t bump_cap = 0, bump_used = 0;
    ~mmc_hip_slotres() {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

static int slot_init(mmc_hip_slotres* sl) {
    /* non-blocking: a slot's launch neither waits for nor holds up work on
     * the null stream or on other slots */
    HIPCHK(hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&sl->e0));
    HIPCHK(hipEventCreate(&sl->e1));
    return 0;
}

/* Builds the tables of one persistent launch over nsim simulations, copies
 * them and enqueues the kernel on the slot's stream; returns without
 * waiting.  Without an arena the fields are allocated here and reduced to
 * FP64 by slot_finish; with an arena they accumulate at the caller's
 * offsets and stay there.  res is only read (record windows, detector
 * buffers) and gets each simulation's reclen. */
static int32_t slot_enqueue(mmc_hip_slotres* sl, mmc_hip_mesh_handle m, const mmc_hip_sim* sims, int32_t nsim,
                            const mmc_hip_tune* tune, mmc_hip_result* res, mmc_hip_arenares* arena,
                            const uint64_t* arena_off) {
    if (sl->busy) return set_err(-1, "launch slot busy: wait for its launch first");
    sl->t_start = std::chrono::steady_clock::now();
    sl->bump_used = 0;
    /* a per-launch device buffer: pooled, or a window of the scratch array */
    auto need = [sl](Pooled& b, size_t bytes) -> int {
        if (!sl->bump_base) return b.ensure(bytes);
        const size_t off = (sl->bump_used + 255) & ~(size_t)255;
        if (off + bytes > sl->bump_cap)
            return set_err(-1, "allocation-free launch: its buffers need more than the %zu bytes of scratch",
                           sl->bump_cap);
        b.p = sl->bump_base + off;
        b.cap = bytes;
        b.fixed = true;
        sl->bump_used = off + std::max<size_t>(bytes, 1);
        return 0;
    };
    /* own fields left by a launch that failed before it was enqueued */
    if (sl->field.p) (void)hipFree(sl->field.p);
    if (sl->lo.p) (void)hipFree(sl->lo.p);
    sl->field.p = sl->lo.p = nullptr;
    if (!m || !sims || nsim < 1 || !res) return set_err(-1, "null argument");
    {
        int rc = ensure_libm_tables();
        if (rc) return rc;
    }
    for (int i = 0; i < nsim; i++) {
        int rc = check_sim(m, sims + i, sims[0].method);
        if (rc) return set_err(rc, "simulation %d: %s", i, std::string(g_err).c_str());
    }
    const bool rf = rf_of(&sims[0]);
    const int tracer = sims[0].method == MMC_RT_PLUCKER ? kTrNodal
                       : sims[0].method == MMC_RT_BLBADOUELGRID ? (rf ? kTrGridRF : kTrGrid)
                       : sims[0].basisorder == 1 ? kTrElemNodal
                                                 : (rf ? kTrElemRF : kTrElem);
    /* media rows 0..prop, then the detector label's row: a copy of the
     * background (src/mmc_mesh.c:518-520) */
    const int nmed = m->prop + 1 + m->isextdet;
    int block = (tune && tune->block > 0) ? tune->block : kBlock;
    if (block % 64 || block > kBlock) return set_err(-1, "block must be a multiple of 64 and <= %d", kBlock);
    const bool comp = arena ? (arena->lo != nullptr) : (tune && tune->compensated);

    /* ---- simulation tables (kept in the slot until the launch ends) ---- */
    std::vector<SimHot>& hot = sl->h_hot;
    std::vector<SimCold>& cold = sl->h_cold;
    std::vector<float4>& media = sl->h_media;
    std::vector<int64_t>& cum = sl->cum;
    std::vector<float4>& dets = sl->h_dets;
    std::vector<float>& pats = sl->h_pats;
    std::vector<uint64_t>& seeds = sl->h_seeds;
    std::vector<float>& rws = sl->h_rw;
    std::vector<float>& rts = sl->h_rt;
    std::vector<SimGrid>& grids = sl->h_grid;
    const bool gridout = tracer == kTrGrid || tracer == kTrGridRF;
    grids.assign(gridout ? nsim : 0, SimGrid());
    seeds.clear();
    rws.clear();
    rts.clear();
    hot.assign(nsim, SimHot());
    cold.assign(nsim, SimCold());
    media.assign((size_t)nsim * nmed, float4());
    cum.assign(nsim + 1, 0);
    dets.clear();
    pats.clear();
    uint64_t fieldlen = 0;
    std::vector<uint64_t> fieldoff(nsim);
    int savedet = 0, maxrec = 0, maxsrc = 1;
    int64_t maxdet = 0;
    for (int i = 0; i < nsim; i++) {
        const mmc_hip_sim& s = sims[i];
        SimHot& h = hot[i];
        SimCold& c = cold[i];
        memset(&h, 0, sizeof(h));
        memset(&c, 0, sizeof(c));
        h.tstart = s.tstart;
        h.tend = s.tend;
        h.rtstep = 1.f / s.tstep;
        h.minenergy = s.minenergy;
        h.roulettesize = s.roulettesize;
        h.maxgate = s.maxgate;
        /* roulette only for single-gate runs (src/mmc_raytrace.c:2033) */
        float span = (s.tend - s.tstart) * h.rtstep;
        const int rmode = s.outputtype == MMC_OUT_JACOBIAN ? kReplayJ
                          : s.outputtype == MMC_OUT_WL ? kReplayWL
                          : s.outputtype == MMC_OUT_WP ? kReplayWP : kReplayNone;
        h.flags = ((s.minenergy > 0.f && span <= 1.f && rmode == kReplayNone) ? kFlagRoulette : 0) |
                  ((s.outputtype == MMC_OUT_ENERGY || rmode == kReplayWL || rmode == kReplayWP) ? kFlagEnergy : 0) |
                  (rmode << kReplayShift) |
                  ((s.issavedet && (s.detnum > 0 || m->isextdet)) ? kFlagSaveDet : 0) |
                  (s.voidtime ? kFlagVoidTime : 0) | ((s.isreflect & 3) << kReflShift) |
                  ((s.isspecular & 3) << kSpecShift) | ((s.debuglevel & MMC_DL_TRAJ) ? kFlagTraj : 0) |
                  (mom_of(&s, m) ? kFlagMom : 0);
        h.nout = s.nout;
        h.srcnum = slots_of(&s);
        maxsrc = std::max(maxsrc, h.srcnum);
        if (rf_of(&s) != rf_of(&sims[0]) || (rf_of(&s) && s.omega != sims[0].omega))
            return set_err(-1, "simulation %d: all simulations of a batch must share omega", i);
        if (s.method != MMC_RT_BLBADOUELGRID && s.basisorder != sims[0].basisorder)
            return set_err(-1, "simulation %d: all simulations of a batch must share the output basis", i);
        const uint64_t flen = (uint64_t)datalen_of(m, &s) * s.maxgate * (rf_of(&s) ? 2 : 1);
        if (arena) {
            if (arena_off[i] + flen > arena->n)
                return set_err(-1, "simulation %d: field [%llu, +%llu) outside the arena", i,
                               (unsigned long long)arena_off[i], (unsigned long long)flen);
            fieldoff[i] = arena_off[i];
        } else {
            fieldoff[i] = fieldlen;
            fieldlen += flen;
        }
        h.fieldoff = fieldoff[i];
        c.srcpos = make_float4(s.srcpos[0], s.srcpos[1], s.srcpos[2], s.srcpos[3]);
        c.srcdir = make_float4(s.srcdir[0], s.srcdir[1], s.srcdir[2], s.srcdir[3]);
        c.srcparam1 = make_float4(s.srcparam1[0], s.srcparam1[1], s.srcparam1[2], s.srcparam1[3]);
        c.srcparam2 = make_float4(s.srcparam2[0], s.srcparam2[1], s.srcparam2[2], s.srcparam2[3]);
        c.bary0 = make_float4(s.bary0[0], s.bary0[1], s.bary0[2], s.bary0[3]);
        c.tstep = s.tstep;
        c.nout = s.nout;
        c.minenergy = s.minenergy;
        c.roulettesize = s.roulettesize;
        c.e0 = s.e0;
        c.srctype = s.srctype;
        c.srcnum = slots_of(&s);
        c.isreflect = s.isreflect;
        c.isspecular = s.isspecular;
        c.voidtime = s.voidtime;
        c.outputtype = s.outputtype;
        /* wide-field detectors replace the point detectors (src/mmc_mesh.c:402-406) */
        c.issavedet = (s.issavedet && (s.detnum > 0 || m->isextdet));
        c.issaveexit = (c.issavedet && s.issaveexit > 0) ? 1 : 0;
        c.detnum = (c.issavedet && !m->isextdet) ? s.detnum : 0;
        c.detbase = (int)dets.size();
        for (int d = 0; d < c.detnum; d++)
            dets.push_back(make_float4(s.detpos[4 * d], s.detpos[4 * d + 1], s.detpos[4 * d + 2], s.detpos[4 * d + 3]));
        if (s.srctype == MMC_SRC_PATTERN) {
            c.patnx = (int)s.srcparam1[3];
            c.patny = (int)s.srcparam2[3];
            c.patbase = (int)pats.size();
            h.patbase = c.patbase;
            pats.insert(pats.end(), s.srcpattern, s.srcpattern + (size_t)c.patnx * c.patny * c.srcnum);
        } else if (s.srctype == MMC_SRC_PATTERN3D) {
            c.patnx = (int)s.srcparam1[0];
            c.patny = (int)s.srcparam1[1];
            c.patbase = (int)pats.size();
            h.patbase = c.patbase;
            pats.insert(pats.end(), s.srcpattern, s.srcpattern + (size_t)c.patnx * c.patny * (int)s.srcparam1[2]);
        }
        c.nphoton = s.nphoton;
        c.photon_first = s.photon_offset;
        c.seed = s.seed;
        c.fieldoff = fieldoff[i];
        c.replaybase = -1;
        if (s.photonseed) {
            c.replaybase = (int64_t)(seeds.size() / 2);
            const uint64_t* sp = s.photonseed + 2 * (size_t)s.photon_offset;
            seeds.insert(seeds.end(), sp, sp + 2 * (size_t)s.nphoton);
            for (int64_t k = 0; k < s.nphoton; k++) {
                rws.push_back(s.replayweight ? s.replayweight[s.photon_offset + k] : 1.f);
                rts.push_back(s.replaytime ? s.replaytime[s.photon_offset + k] : 0.f);
            }
        }
        if (gridout) grids[i] = grid_of(m, &s);
        const float* md = s.med ? s.med : m->med.data();
        for (int k = 0; k <= m->prop; k++)
            media[(size_t)i * nmed + k] = make_float4(md[4 * k], md[4 * k + 1], md[4 * k + 2], md[4 * k + 3]);
        if (m->isextdet) media[(size_t)i * nmed + m->prop + 1] = media[(size_t)i * nmed];
        cum[i + 1] = cum[i] + s.nphoton;
        /* record: detector id, scattering counts, partial paths, momentum
         * transfer (ismomentum), exit position and direction (issaveexit),
         * launch weight; the device rows use the longest record's stride */
        res[i].reclen = c.issavedet ? 1 + (2 + mom_of(&sims[i], m)) * m->prop + 6 * c.issaveexit + 1 : 0;
        if (c.issavedet) {
            savedet = 1;
            maxrec = std::max(maxrec, res[i].reclen);
            maxdet += s.maxdetphoton > 0 ? s.maxdetphoton : 0;
        }
    }
    if (arena) {
        /* the simulations of a launch write disjoint arena regions */
        std::vector<std::pair<uint64_t, uint64_t>> reg(nsim);
        for (int i = 0; i < nsim; i++)
            reg[i] = {fieldoff[i], (uint64_t)datalen_of(m, sims + i) * sims[i].maxgate * (rf_of(sims + i) ? 2 : 1)};
        std::sort(reg.begin(), reg.end());
        for (int i = 1; i < nsim; i++)
            if (reg[i - 1].second > 0 && reg[i].second > 0 && reg[i - 1].first + reg[i - 1].second > reg[i].first)
                return set_err(-1, "arena regions of the launch overlap at offset %llu",
                               (unsigned long long)reg[i].first);
    }
    const int64_t total = cum[nsim];
    const int reclen = maxrec;

    /* ---- device buffers ---- */
    int nrep = 1;
    if (arena) {
        nrep = arena->nrep;
    } else if (tune && tune->replicas > 0) {
        nrep = tune->replicas <= 8 ? tune->replicas : std::min(tune->replicas / 8 * 8, kMaxReplicas);
    } else {
        nrep = auto_replicas((uint64_t)fieldlen * sizeof(float) * (comp ? 2 : 1));
    }
    int rc;
    if (sl->bump_base && !arena) return set_err(-1, "allocation-free launch: needs an arena for its fields");
    if ((rc = need(sl->d_hot, sizeof(SimHot) * nsim)) || (rc = need(sl->d_cold, sizeof(SimCold) * nsim)) ||
        (rc = need(sl->d_med, sizeof(float4) * media.size())) || (rc = need(sl->d_cum, sizeof(int64_t) * cum.size())) ||
        (rc = need(sl->d_det, sizeof(float4) * dets.size())) || (rc = need(sl->d_pat, sizeof(float) * pats.size())) ||
        (rc = need(sl->d_acc, sizeof(double) * nsim * kAccStride)) ||
        (rc = need(sl->d_detcnt, sizeof(unsigned long long) * (nsim + 2))) ||
        (rc = need(sl->d_queue, sizeof(unsigned long long))))
        return rc;
    if (!seeds.empty() &&
        ((rc = need(sl->d_seeds, sizeof(uint64_t) * seeds.size())) || (rc = need(sl->d_rw, sizeof(float) * rws.size())) ||
         (rc = need(sl->d_rt, sizeof(float) * rts.size()))))
        return rc;
    if (!grids.empty() && (rc = need(sl->d_grid, sizeof(SimGrid) * grids.size()))) return rc;
    if (!arena) {
        if ((rc = dalloc((float**)&sl->field.p, (size_t)fieldlen * nrep, nullptr))) return rc;
        if (comp && (rc = dalloc((float**)&sl->lo.p, (size_t)fieldlen * nrep, nullptr))) return rc;
    }
    if (savedet && maxdet > 0) {
        if ((rc = need(sl->d_detdata, sizeof(float) * (size_t)maxdet * reclen)) ||
            (rc = need(sl->d_detph, sizeof(int64_t) * (size_t)maxdet)))
            return rc;
    }
    /* per-photon records: one window, taken from the first simulation's request */
    int64_t recfirst = 0, reccount = 0;
    for (int i = 0; i < nsim; i++)
        if (res[i].photons && res[i].reccount > 0) {
            recfirst = cum[i] + res[i].recfirst;
            reccount = std::min(res[i].reccount, sims[i].nphoton - res[i].recfirst);
            break;
        }
    if (reccount > 0 && (rc = need(sl->d_rec, sizeof(mmc_hip_photon) * (size_t)reccount))) return rc;
    /* trajectory records: one buffer per launch, returned through the first
     * result that provides one */
    int trajres = -1;
    int64_t maxtraj = 0;
    for (int i = 0; i < nsim; i++)
        if ((sims[i].debuglevel & MMC_DL_TRAJ) && res[i].traj && res[i].maxtraj > 0) {
            trajres = i;
            maxtraj = res[i].maxtraj;
            break;
        }
    if (maxtraj > 0 && (rc = need(sl->d_traj, sizeof(float) * MMC_TRAJ_RECLEN * (size_t)maxtraj))) return rc;

    hipStream_t st = sl->st;
    HIPCHK(hipMemcpyAsync(sl->d_hot.p, hot.data(), sizeof(SimHot) * nsim, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(sl->d_cold.p, cold.data(), sizeof(SimCold) * nsim, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(sl->d_med.p, media.data(), sizeof(float4) * media.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(sl->d_cum.p, cum.data(), sizeof(int64_t) * cum.size(), hipMemcpyHostToDevice, st));
    if (!dets.empty())
        HIPCHK(hipMemcpyAsync(sl->d_det.p, dets.data(), sizeof(float4) * dets.size(), hipMemcpyHostToDevice, st));
    if (!pats.empty())
        HIPCHK(hipMemcpyAsync(sl->d_pat.p, pats.data(), sizeof(float) * pats.size(), hipMemcpyHostToDevice, st));
    if (!grids.empty())
        HIPCHK(hipMemcpyAsync(sl->d_grid.p, grids.data(), sizeof(SimGrid) * grids.size(), hipMemcpyHostToDevice, st));
    if (!seeds.empty()) {
        HIPCHK(hipMemcpyAsync(sl->d_seeds.p, seeds.data(), sizeof(uint64_t) * seeds.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(sl->d_rw.p, rws.data(), sizeof(float) * rws.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(sl->d_rt.p, rts.data(), sizeof(float) * rts.size(), hipMemcpyHostToDevice, st));
    }
    if (!arena) {
        HIPCHK(hipMemsetAsync(sl->field.p, 0, sizeof(float) * (fieldlen ? fieldlen * nrep : 1), st));
        if (sl->lo.p) HIPCHK(hipMemsetAsync(sl->lo.p, 0, sizeof(float) * (fieldlen ? fieldlen * nrep : 1), st));
    }
    /* energy tallies, queue and counters are cleared by fill kernels (no
     * memset nodes when a caller captures the launch in a graph) */
    hipLaunchKernelGGL(fill_zero_u64, dim3((nsim * kAccStride + 63) / 64), dim3(64), 0, st,
                       (unsigned long long*)sl->d_acc.p, nsim * kAccStride);
    hipLaunchKernelGGL(fill_zero_u64, dim3(1), dim3(64), 0, st, (unsigned long long*)sl->d_queue.p, 1);
    hipLaunchKernelGGL(fill_zero_u64, dim3((nsim + 65) / 64), dim3(64), 0, st,
                       (unsigned long long*)sl->d_detcnt.p, nsim + 2);

    KArgs a;
    memset(&a, 0, sizeof(a));
    a.planes = m->planes;
    a.facenb = m->facenb;
    a.lblshift = m->lblshift;
    a.elem = m->elem;
    a.type = m->type;
    a.node = m->node;
    a.enode = m->enode;
    a.srcelem = m->srcelem;
    a.srcelemlen = m->srcelemlen;
    a.loc_start = m->loc_start;
    a.loc_elem = m->loc_elem;
    a.loc_min = m->loc_min;
    a.loc_inv = m->loc_inv;
    a.loc_dim = m->loc_dim;
    a.nn = m->nn;
    a.ne = m->ne;
    a.prop = m->prop;
    a.isextdet = m->isextdet;
    a.nsim = nsim;
    a.hot = (const SimHot*)sl->d_hot.p;
    a.cold = (const SimCold*)sl->d_cold.p;
    a.media = (const float4*)sl->d_med.p;
    a.cum = (const int64_t*)sl->d_cum.p;
    a.detpos = (const float4*)sl->d_det.p;
    a.pattern = (const float*)sl->d_pat.p;
    a.field = arena ? arena->field : (float*)sl->field.p;
    a.field_lo = arena ? arena->lo : (float*)sl->lo.p;
    a.nrep = nrep;
    a.repstride = arena ? arena->n : fieldlen;
    a.simacc = (double*)sl->d_acc.p;
    a.detdata = (savedet && maxdet > 0) ? (float*)sl->d_detdata.p : nullptr;
    a.detphoton = (savedet && maxdet > 0) ? (int64_t*)sl->d_detph.p : nullptr;
    a.detcount = (unsigned long long*)sl->d_detcnt.p;
    a.maxdet = (savedet && maxdet > 0) ? maxdet : 0;
    a.reclen = reclen;
    a.savedet = savedet;
    a.rec = reccount > 0 ? (mmc_hip_photon*)sl->d_rec.p : nullptr;
    a.recfirst = recfirst;
    a.reccount = reccount;
    a.queue = (unsigned long long*)sl->d_queue.p;
    a.total = total;
    a.chunk = (tune && tune->chunk > 0) ? tune->chunk : 1;
    a.maxsteps = MMC_MAX_STEPS;
    a.maxsrc = maxsrc;
    a.seeds = seeds.empty() ? nullptr : (const uint64_t*)sl->d_seeds.p;
    a.rweight = seeds.empty() ? nullptr : (const float*)sl->d_rw.p;
    a.rtime = seeds.empty() ? nullptr : (const float*)sl->d_rt.p;
    a.grid = grids.empty() ? nullptr : (const SimGrid*)sl->d_grid.p;
    if (gridout && !a.grid) return set_err(-1, "internal: dual-grid launch without its grid table");
    a.traj = maxtraj > 0 ? (float*)sl->d_traj.p : nullptr;
    a.maxtraj = (uint64_t)maxtraj;
    a.omega = rf ? sims[0].omega : 0.f;
    int anymom = 0;
    for (int i = 0; i < nsim; i++) anymom |= mom_of(&sims[i], m);
    a.mom = anymom;

    const int nslots = kLaneSlots + (2 + anymom) * m->prop + 3 + (maxsrc > 1 ? 3 * maxsrc : 0) + m->isextdet;
    /* per-simulation tables go to LDS while they take at most 16 KiB, so that
     * large batches keep 4+ workgroups per CU */
    const size_t simtab = (sizeof(SimHot) + sizeof(float4) * (size_t)nmed) * nsim;
    const bool stage = simtab <= 16384;
    /* LDS: the staged tables plus per-lane slots (energy tallies, and with
     * detectors a scattering count, partial path (and momentum transfer) per
     * medium); meshes with many media shrink the workgroup to fit, unless
     * the caller fixed it.  The kernel's static LDS (agreement thresholds)
     * comes on top. */
    const size_t kLdsMax = 160 * 1024 - 2048;
    auto lds_for = [&](int b) {
        size_t v = (stage ? simtab : 0) + sizeof(int64_t) * (nsim + 1) + sizeof(float) * (size_t)nslots * b;
        return (v + 15) & ~(size_t)15;
    };
    while (lds_for(block) > kLdsMax && block > 64 && !(tune && tune->block > 0)) block /= 2;
    if (lds_for(block) > kLdsMax)
        return set_err(-1, "batch of %d simulations needs %zu B of LDS (%d media%s)", nsim, lds_for(block), m->prop,
                       savedet ? " with detector records" : "");

    /* the register budget of the kernel variant follows the requested
     * residency: 4 workgroups per CU by default (128 VGPRs), 5 or more select
     * the variant compiled for 5; large batches read their tables through the
     * caches (STAGE=false) */
    const int want = (tune && tune->blocks_per_cu > 0) ? tune->blocks_per_cu : 0;
    const void* kfn = nullptr;
#define MMC_KFN(T, O, S) (comp ? (const void*)transport<T, true, O, S> : (const void*)transport<T, false, O, S>)
#define MMC_PICK(O, S)                                                                 \
    kfn = tracer == kTrNodal    ? MMC_KFN(kTrNodal, O, S)                               \
          : tracer == kTrElemNodal ? MMC_KFN(kTrElemNodal, O, S)                        \
          : tracer == kTrGrid   ? MMC_KFN(kTrGrid, O, S)                                \
          : tracer == kTrGridRF ? MMC_KFN(kTrGridRF, O, S)                              \
          : tracer == kTrElemRF ? MMC_KFN(kTrElemRF, O, S)                              \
                                : MMC_KFN(kTrElem, O, S);
    /* 3 workgroups per CU leave the Plucker kernel 168 VGPRs and no spills:
     * against 4 per CU, 1.07x with one pattern and 1.27x with 8 shared
     * patterns on the 24.0M-tet shells mesh, 1.09x on the 162k-tet cube.  The
     * Badouel kernels run best at 4 (the 3-per-CU build is 13 % slower on the
     * slab). */
    const bool three = want == 3 || (want == 0 && tracer == kTrNodal);
    if (!stage) {
        MMC_PICK(4, kTabGlobal)
    } else if (want >= 5) {
        MMC_PICK(5, kTabLds)
    } else if (three) {
        MMC_PICK(3, kTabLds)
    } else {
        MMC_PICK(4, kTabLds)
    }
#undef MMC_PICK
#undef MMC_KFN
    int dev = 0, ncu = 0, maxbpc = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&maxbpc, kfn, block, lds_for(block)));
    /* When the per-lane LDS slots (detector records of many media, pattern
     * weights and energies of shared patterns) leave room for fewer than
     * the register-limited number of waves per CU, smaller workgroups pack
     * the LDS more tightly: take the workgroup size with the most resident
     * waves (the larger on ties).  A caller that fixed the workgroup size
     * or the workgroups per CU keeps them. */
    if (!(tune && tune->block > 0) && want == 0) {
        for (int b = block / 2; b >= 64; b /= 2) {
            if (lds_for(b) > kLdsMax) continue;
            int n = 0;
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kfn, b, lds_for(b)));
            if ((int64_t)n * b > (int64_t)maxbpc * block) {
                block = b;
                maxbpc = n;
            }
        }
    }
    const size_t shmem = lds_for(block);
    int bpc = want > 0 ? std::min(want, std::max(maxbpc, 1)) : std::max(maxbpc, 1);
    int grid = ncu * bpc;

    if (sl->e0) HIPCHK(hipEventRecord(sl->e0, st));
    if (total > 0) {
        void* kargs[] = {&a};
        HIPCHK(hipLaunchKernel(kfn, dim3(grid), dim3(block), kargs, shmem, st));
        HIPCHK(hipGetLastError());
    }
    if (sl->e1) HIPCHK(hipEventRecord(sl->e1, st));

    /* state slot_finish needs */
    sl->busy = true;
    sl->m = m;
    sl->arena = arena;
    sl->nsim = nsim;
    sl->nrep = nrep;
    sl->reclen = reclen;
    sl->savedet = savedet;
    sl->comp = comp;
    sl->maxdet = maxdet;
    sl->recfirst = recfirst;
    sl->reccount = reccount;
    sl->trajres = trajres;
    sl->maxtraj = maxtraj;
    sl->fieldlen = fieldlen;
    sl->fieldoff = fieldoff;
    sl->datalen.resize(nsim);
    sl->maxdetph.resize(nsim);
    sl->photon_offset.resize(nsim);
    sl->nphoton.resize(nsim);
    sl->maxgate.resize(nsim);
    sl->nslot.resize(nsim);
    sl->rfmul.resize(nsim);
    for (int i = 0; i < nsim; i++) {
        sl->datalen[i] = datalen_of(m, sims + i);
        sl->maxdetph[i] = sims[i].maxdetphoton;
        sl->photon_offset[i] = sims[i].photon_offset;
        sl->nphoton[i] = sims[i].nphoton;
        sl->maxgate[i] = sims[i].maxgate;
        sl->nslot[i] = slots_of(sims + i);
        sl->rfmul[i] = rf_of(sims + i) ? 2 : 1;
    }
    return 0;
}

/* Waits for the slot's launch and adds its results into res[0..nsim):
 * energy tallies, ray-tet counts, kernel and wall times, detected records,
 * per-photon records and (without an arena) the FP64 field and its sum. */
static int32_t slot_finish(mmc_hip_slotres* sl, mmc_hip_result* res) {
    if (!sl->busy) return set_err(-1, "launch slot idle: nothing to wait for");
    sl->busy = false;
    struct FieldRelease {
        mmc_hip_slotres* s;
        ~FieldRelease() {
            if (s->field.p) (void)hipFree(s->field.p);
            if (s->lo.p) (void)hipFree(s->lo.p);
            s->field.p = s->lo.p = nullptr;
        }
    } fr{sl};
    const int nsim = sl->nsim;
    hipStream_t st = sl->st;
    std::vector<double> acc((size_t)nsim * kAccStride);
    std::vector<unsigned long long> cnt(nsim + 2);
    HIPCHK(hipMemcpyAsync(acc.data(), sl->d_acc.p, sizeof(double) * acc.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cnt.data(), sl->d_detcnt.p, sizeof(unsigned long long) * cnt.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float kms = 0.f;
    if (sl->e0 && sl->e1) {
        HIPCHK(hipEventElapsedTime(&kms, sl->e0, sl->e1));
    } else {
        /* allocation-free launch (no events): the wall time of the launch */
        kms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - sl->t_start).count();
    }
    for (int i = 0; i < nsim; i++) {
        mmc_hip_result& r = res[i];
        const double* A = acc.data() + (size_t)i * kAccStride;
        r.launchweight += A[kAccLaunch];
        r.absorbweight += A[kAccAbsorb];
        r.escapeweight += A[kAccEscape];
        r.roulettegain += A[kAccRGain];
        r.raytet += A[kAccRaytet];
        r.detected += (int64_t)cnt[1 + i];
        r.kernel_ms += kms;
        if (r.srcweight && sl->nslot[i] > 1)
            for (int k = 0; k < sl->nslot[i]; k++) {
                r.srcweight[k] += A[kAccSrcLaunch + k];
                r.srcweight[sl->nslot[i] + k] += A[kAccSrcAbsorb + k];
            }
    }

    /* FP64 final reduction of every simulation's own field on the device; the
     * host copy happens only when the caller passed a field array */
    if (!sl->arena) {
        int dev = 0, ncu = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        /* the FP64 staging buffer exists only when a field goes to the host;
         * otherwise the reduction yields the sums alone */
        uint64_t maxn = 1;
        bool host = false;
        for (int i = 0; i < nsim; i++) {
            maxn = std::max<uint64_t>(maxn, (uint64_t)sl->datalen[i] * sl->maxgate[i] * sl->rfmul[i]);
            host = host || res[i].field || res[i].field_im;
        }
        DevBuf d64, dsum;
        int rc;
        if ((host && (rc = dalloc((double**)&d64.p, (size_t)maxn, nullptr))) ||
            (rc = dalloc((double**)&dsum.p, (size_t)nsim + 1, nullptr)))
            return rc;
        HIPCHK(hipMemsetAsync(dsum.p, 0, sizeof(double) * (nsim + 1), st));
        for (int i = 0; i < nsim; i++) {
            mmc_hip_result& r = res[i];
            /* RF: the imaginary block follows the real one; both are
             * reduced and the sum covers the real part only */
            const uint64_t nre = (uint64_t)sl->datalen[i] * sl->maxgate[i];
            if (nre == 0 || sl->nphoton[i] == 0) continue;
            int rb = (int)std::min<uint64_t>((nre + 255) / 256, (uint64_t)ncu * 16);
            for (int part = 0; part < sl->rfmul[i]; part++) {
                const uint64_t off = sl->fieldoff[i] + (uint64_t)part * nre;
                hipLaunchKernelGGL(field_reduce, dim3(std::max(rb, 1)), dim3(256), 0, st, (float*)sl->field.p + off,
                                   sl->lo.p ? (float*)sl->lo.p + off : nullptr, (double*)d64.p, nre,
                                   (double*)dsum.p + (part ? nsim : i), sl->nrep, (uint64_t)sl->fieldlen);
                HIPCHK(hipGetLastError());
                double* dst = part ? r.field_im : r.field;
                if (dst) {
                    std::unique_ptr<double[]> hbuf(new (std::nothrow) double[nre]);
                    if (!hbuf) return set_err(-2, "host staging of %llu field values failed", (unsigned long long)nre);
                    HIPCHK(hipMemcpyAsync(hbuf.get(), d64.p, sizeof(double) * nre, hipMemcpyDeviceToHost, st));
                    HIPCHK(hipStreamSynchronize(st));
                    add_into(dst, hbuf.get(), nre);
                }
            }
        }
        std::vector<double> sums(nsim, 0.0);
        HIPCHK(hipMemcpyAsync(sums.data(), dsum.p, sizeof(double) * nsim, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int i = 0; i < nsim; i++) res[i].fieldsum += sums[i];
    }
    /* detected photons: saved records are split per simulation by photon id range */
    if (sl->savedet && sl->maxdet > 0 && cnt[0] > 0) {
        const int reclen = sl->reclen;
        int64_t nsaved = std::min<int64_t>((int64_t)cnt[0], sl->maxdet);
        std::vector<float> dd((size_t)nsaved * reclen);
        std::vector<int64_t> dp((size_t)nsaved);
        HIPCHK(hipMemcpyAsync(dd.data(), sl->d_detdata.p, sizeof(float) * dd.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(dp.data(), sl->d_detph.p, sizeof(int64_t) * dp.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::vector<int64_t> fill(nsim, 0);
        for (int64_t k = 0; k < nsaved; k++) {
            /* records carry the global photon index; cum[] maps it to its simulation */
            int s = (int)(std::upper_bound(sl->cum.begin(), sl->cum.end(), dp[k]) - sl->cum.begin()) - 1;
            if (s < 0 || s >= nsim) continue;
            mmc_hip_result& r = res[s];
            if (!r.detdata || fill[s] >= sl->maxdetph[s]) continue;
            memcpy(r.detdata + (size_t)fill[s] * r.reclen, dd.data() + (size_t)k * reclen, sizeof(float) * r.reclen);
            if (r.detphoton) r.detphoton[fill[s]] = sl->photon_offset[s] + (dp[k] - sl->cum[s]);
            fill[s]++;
        }
    }
    if (sl->trajres >= 0) {
        mmc_hip_result& r = res[sl->trajres];
        const int64_t made = (int64_t)cnt[nsim + 1];
        const int64_t keep = std::min<int64_t>(made, sl->maxtraj);
        if (keep > 0)
            HIPCHK(hipMemcpyAsync(r.traj, sl->d_traj.p, sizeof(float) * MMC_TRAJ_RECLEN * (size_t)keep,
                                  hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        r.ntraj += made;
    }
    if (sl->reccount > 0) {
        for (int i = 0; i < nsim; i++)
            if (res[i].photons && res[i].reccount > 0) {
                HIPCHK(hipMemcpyAsync(res[i].photons, sl->d_rec.p, sizeof(mmc_hip_photon) * sl->reccount,
                                      hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                break;
            }
    }
    double tms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - sl->t_start).count();
    for (int i = 0; i < nsim; i++) res[i].total_ms += tms;
    return 0;
}

/* Synchronous launch on a temporary slot. */
static int32_t launch_batch(mmc_hip_mesh_handle m, const mmc_hip_sim* sims, int32_t nsim,
                            const mmc_hip_tune* tune, mmc_hip_result* res, mmc_hip_arenares* arena,
                            const uint64_t* arena_off) {
    mmc_hip_slotres sl;
    int rc = slot_init(&sl);
    if (!rc) rc = slot_enqueue(&sl, m, sims, nsim, tune, res, arena, arena_off);
    if (!rc) rc = slot_finish(&sl, res);
    return rc;
}

extern "C" int32_t mmc_hip_run_batch(mmc_hip_mesh_handle m, const mmc_hip_sim* sims, int32_t nsim,
                                     const mmc_hip_tune* tune, mmc_hip_result* res) {
    return launch_batch(m, sims, nsim, tune, res, nullptr, nullptr);
}

extern "C" int32_t mmc_hip_arena_create(uint64_t nfloat, int32_t compensated, int32_t replicas,
                                        mmc_hip_arena_handle* out) {
    if (!out || nfloat == 0) return set_err(-1, "bad arena request");
    if (replicas < 1 || replicas > 8) return set_err(-1, "arena replicas must be 1..8");
    mmc_hip_arenares* a = new mmc_hip_arenares();
    a->n = nfloat;
    a->nrep = replicas;
    const size_t total = (size_t)nfloat * replicas;
    int64_t dummy = 0;
    int rc = 0;
    if (hipStreamCreateWithFlags(&a->st, hipStreamNonBlocking) != hipSuccess) rc = set_err(-3, "arena stream creation failed");
    if (!rc) rc = dalloc(&a->field, total, &dummy);
    if (!rc && compensated) rc = dalloc(&a->lo, total, &dummy);
    if (!rc) {
        hipError_t e = hipMemset(a->field, 0, sizeof(float) * total);
        if (!e && a->lo) e = hipMemset(a->lo, 0, sizeof(float) * total);
        if (e) rc = set_err(-3, "arena clear failed: %s", hipGetErrorString(e));
    }
    if (rc) {
        if (a->field) (void)hipFree(a->field);
        if (a->lo) (void)hipFree(a->lo);
        if (a->st) (void)hipStreamDestroy(a->st);
        delete a;
        return rc;
    }
    *out = a;
    return 0;
}

extern "C" int32_t mmc_hip_arena_release(mmc_hip_arena_handle a) {
    if (!a) return 0;
    if (a->st) {
        (void)hipStreamSynchronize(a->st);
        (void)hipStreamDestroy(a->st);
    }
    if (a->field) (void)hipFree(a->field);
    if (a->lo) (void)hipFree(a->lo);
    delete a;
    return 0;
}

extern "C" int32_t mmc_hip_run_chunks(mmc_hip_mesh_handle m, const mmc_hip_sim* sims, int32_t nsim,
                                      mmc_hip_arena_handle arena, const uint64_t* fieldoff,
                                      const mmc_hip_tune* tune, mmc_hip_result* res) {
    if (!arena || !fieldoff) return set_err(-1, "null arena");
    return launch_batch(m, sims, nsim, tune, res, arena, fieldoff);
}

/* ---- allocation-free launch ------------------------------------------------
 * mmc_hip_run_chunks without device allocation: the per-launch buffers
 * (simulation tables, energy tallies, counters, work queue, and detector,
 * photon, trajectory and replay buffers when asked for) are windows of
 * `scratch`, a file-scope device array of MMC_NOALLOC_SCRATCH bytes of which
 * each device holds its own copy until the process ends.  Its address on
 * each device is looked up once, with that device current, and only that
 * address is handed to copies and kernels.  Launches go to the null stream;
 * calls on one device are serialized here, and launches of other callers on
 * the same device must not overlap them (they would share the array). */
static __device__ __align__(256) unsigned char scratch[MMC_NOALLOC_SCRATCH];
static void* g_scratch_ptr[64];
static mmc_hip_slotres g_noalloc_slot[64];
static std::mutex g_noalloc_mu[64];

extern "C" int32_t mmc_hip_run_chunks_noalloc(mmc_hip_mesh_handle m, const mmc_hip_sim* sims, int32_t nsim,
                                              mmc_hip_arena_handle arena, const uint64_t* fieldoff,
                                              const mmc_hip_tune* tune, mmc_hip_result* res) {
    if (!arena || !fieldoff) return set_err(-1, "null arena");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_err(-1, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu[dev]);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_err(-1, "scratch: device address lookup failed (%s)", hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    mmc_hip_slotres* sl = &g_noalloc_slot[dev];
    sl->bump_base = (unsigned char*)g_scratch_ptr[dev];
    sl->bump_cap = MMC_NOALLOC_SCRATCH;
    int32_t rc = slot_enqueue(sl, m, sims, nsim, tune, res, arena, fieldoff);
    if (!rc) rc = slot_finish(sl, res);
    return rc;
}

extern "C" int32_t mmc_hip_slot_create(mmc_hip_slot_handle* out) {
    if (!out) return set_err(-1, "null argument");
    mmc_hip_slotres* s = new mmc_hip_slotres();
    int rc = slot_init(s);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return 0;
}

extern "C" int32_t mmc_hip_slot_release(mmc_hip_slot_handle s) {
    delete s;
    return 0;
}

extern "C" int32_t mmc_hip_run_chunks_async(mmc_hip_mesh_handle m, const mmc_hip_sim* sims, int32_t nsim,
                                            mmc_hip_arena_handle arena, const uint64_t* fieldoff,
                                            const mmc_hip_tune* tune, mmc_hip_result* res,
                                            mmc_hip_slot_handle slot) {
    if (!arena || !fieldoff || !slot) return set_err(-1, "null arena or slot");
    return slot_enqueue(slot, m, sims, nsim, tune, res, arena, fieldoff);
}

extern "C" int32_t mmc_hip_slot_query(mmc_hip_slot_handle slot) {
    if (!slot) return set_err(-1, "null slot");
    if (!slot->busy) return 1;
    hipError_t e = hipEventQuery(slot->e1);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return set_err(-100 - (int)e, "slot query failed: %s", hipGetErrorString(e));
}

extern "C" int32_t mmc_hip_slot_wait(mmc_hip_slot_handle slot, mmc_hip_result* res) {
    if (!slot || !res) return set_err(-1, "null argument");
    return slot_finish(slot, res);
}

extern "C" int32_t mmc_hip_arena_collect(mmc_hip_arena_handle a, uint64_t off, uint64_t n, double* host,
                                         double* sum) {
    if (!a || off + n > a->n) return set_err(-1, "collect outside the arena");
    if (n == 0) {
        if (sum) *sum = 0.0;
        return 0;
    }
    int rc;
    if ((rc = a->d64.ensure(sizeof(double) * n)) || (rc = a->dsum.ensure(sizeof(double)))) return rc;
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipMemsetAsync(a->dsum.p, 0, sizeof(double), a->st));
    int rb = (int)std::min<uint64_t>((n + 255) / 256, (uint64_t)std::max(ncu, 1) * 16);
    hipLaunchKernelGGL(field_reduce, dim3(std::max(rb, 1)), dim3(256), 0, a->st, a->field + off,
                       a->lo ? a->lo + off : nullptr, (double*)a->d64.p, n, (double*)a->dsum.p, a->nrep, a->n);
    HIPCHK(hipGetLastError());
    std::unique_ptr<double[]> hbuf;
    double s = 0.0;
    if (host) {
        hbuf.reset(new (std::nothrow) double[n]);
        if (!hbuf) return set_err(-2, "host staging of %llu field values failed", (unsigned long long)n);
        HIPCHK(hipMemcpyAsync(hbuf.get(), a->d64.p, sizeof(double) * n, hipMemcpyDeviceToHost, a->st));
    }
    HIPCHK(hipMemcpyAsync(&s, a->dsum.p, sizeof(double), hipMemcpyDeviceToHost, a->st));
    HIPCHK(hipStreamSynchronize(a->st));
    if (host) add_into(host, hbuf.get(), n);
    if (sum) *sum = s;
    return 0;
}

extern "C" int32_t mmc_hip_run(const mmc_hip_mesh* mesh, const mmc_hip_sim* sim, const mmc_hip_tune* tune,
                               mmc_hip_result* res) {
    auto t0 = std::chrono::steady_clock::now();
    mmc_hip_mesh_handle h = nullptr;
    int rc = mmc_hip_mesh_upload(mesh, &h);
    if (rc) return rc;
    double before = res->total_ms;
    rc = mmc_hip_run_batch(h, sim, 1, tune, res);
    mmc_hip_mesh_release(h);
    if (!rc) res->total_ms = before + std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}
