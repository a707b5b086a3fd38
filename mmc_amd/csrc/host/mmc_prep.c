/*
 * mmc_prep.c -- host-side mesh preparation of the MI355X photon tracer.
 *
 * Restates the reference's mmc_prep() chain (src/mmc_host.c:136-165).  The
 * floating-point expressions that feed the transport (volumes, face planes,
 * barycentric coordinates) keep the reference's operation order, because the
 * device kernels are compared photon by photon against the reference
 * semantics; compile with -ffp-contract=off.
 */
#include "mmc_hip_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* face j of an element is spanned by nodes kOut[j] (outward, counter-clockwise),
 * src/mmc_mesh.c:59; kFaceMap[j] is the node opposite to face j */
static const int kOut[4][3] = {{0, 3, 1}, {3, 2, 1}, {0, 2, 3}, {0, 1, 2}};
static const int kFaceMap[4] = {2, 0, 1, 3};
/* facenb column j <-> plane index kIFaceOrder[j] (src/mmc_mesh.c:103) */
static const int kIFaceOrder[4] = {3, 0, 2, 1};
/* nodes of facenb column j (src/mmc_highorder.cpp:50) */
static const int kNbFace[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

typedef struct { float x, y, z; } f3;

static inline f3 at(const float* node, int id1) {
    const float* p = node + 3 * (size_t)(id1 - 1);
    f3 r = {p[0], p[1], p[2]};
    return r;
}
static inline f3 sub(f3 a, f3 b) { f3 r = {b.x - a.x, b.y - a.y, b.z - a.z}; return r; } /* b - a */
static inline f3 crs(f3 a, f3 b) {
    f3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static inline float dt(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

void mmc_host_volume(int32_t nn, const float* node, int32_t ne, int32_t* elem,
                     const int32_t* type, float* evol, float* nvol) {
    memset(nvol, 0, sizeof(float) * (size_t)nn);
    for (int32_t i = 0; i < ne; i++) {
        int32_t* e = elem + 4 * (size_t)i;
        f3 a = at(node, e[0]), b = at(node, e[1]), c = at(node, e[2]), d = at(node, e[3]);
        /* cofactor expansion in the reference's grouping (src/mmc_mesh.c:918-927) */
        float cx = c.x - d.x, cy = c.y - d.y, cz = c.z - d.z;
        float m_yz = c.y * d.z - c.z * d.y;
        float m_xz = c.x * d.z - c.z * d.x;
        float m_xy = c.x * d.y - c.y * d.x;
        float v = b.x * m_yz - b.y * m_xz + b.z * m_xy;
        v += -a.x * (m_yz + b.y * cz - b.z * cy);
        v += +a.y * (m_xz + b.x * cz - b.z * cx);
        v += -a.z * (m_xy + b.x * cy - b.y * cx);
        v = -v;
        if (v < 0.f) {
            int32_t t = e[3];
            e[3] = e[2];
            e[2] = t;
            v = -v;
        }
        v *= (1.f / 6.f);
        evol[i] = v;
        if (type[i] == 0) continue;
        for (int j = 0; j < 4; j++) nvol[e[j] - 1] += v * 0.25f;
    }
}

/* Node -> element incidence (CSR), then for every face the neighbour is the
 * other element incident to all three face nodes. */
int32_t mmc_host_facenb(int32_t nn, int32_t ne, const int32_t* elem, int32_t* facenb) {
    int64_t* start = (int64_t*)calloc((size_t)nn + 1, sizeof(int64_t));
    int32_t* inc = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)ne);
    if (!start || !inc) {
        free(start);
        free(inc);
        return -1;
    }
    for (size_t k = 0; k < 4 * (size_t)ne; k++) {
        int32_t n = elem[k];
        if (n < 1 || n > nn) {
            free(start);
            free(inc);
            return -2;
        }
        start[n]++;
    }
    for (int32_t n = 0; n < nn; n++) start[n + 1] += start[n];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)nn);
    memcpy(fill, start, sizeof(int64_t) * (size_t)nn);
    for (int32_t i = 0; i < ne; i++)
        for (int j = 0; j < 4; j++) inc[fill[elem[4 * (size_t)i + j] - 1]++] = i;
    free(fill);

    for (int32_t i = 0; i < ne; i++) {
        const int32_t* e = elem + 4 * (size_t)i;
        for (int j = 0; j < 4; j++) {
            int32_t a = e[kNbFace[j][0]], b = e[kNbFace[j][1]], c = e[kNbFace[j][2]];
            int32_t nb = 0;
            for (int64_t k = start[a - 1]; k < start[a]; k++) {
                int32_t o = inc[k];
                if (o == i) continue;
                const int32_t* f = elem + 4 * (size_t)o;
                int hb = (f[0] == b) | (f[1] == b) | (f[2] == b) | (f[3] == b);
                int hc = (f[0] == c) | (f[1] == c) | (f[2] == c) | (f[3] == c);
                if (hb && hc) {
                    nb = o + 1;
                    break;
                }
            }
            facenb[4 * (size_t)i + j] = nb;
        }
    }
    free(start);
    free(inc);
    return 0;
}

void mmc_host_planes(int32_t ne, const float* node, const int32_t* elem, float* planes) {
    for (int32_t i = 0; i < ne; i++) {
        const int32_t* e = elem + 4 * (size_t)i;
        float* P = planes + 16 * (size_t)i;
        for (int j = 0; j < 4; j++) {
            f3 A = at(node, e[kOut[j][0]]), B = at(node, e[kOut[j][1]]), C = at(node, e[kOut[j][2]]);
            f3 N = crs(sub(A, B), sub(A, C));
            /* 1.f/sqrt in double, rounded to float (src/mmc_mesh.c:1589) */
            float inv = (float)(1.0 / sqrt((double)dt(N, N)));
            N.x = inv * N.x;
            N.y = inv * N.y;
            N.z = inv * N.z;
            P[j] = N.x;
            P[4 + j] = N.y;
            P[8 + j] = N.z;
            P[12 + j] = dt(N, A);
        }
    }
}

int32_t mmc_host_barycentric(const float* node, const int32_t* elem, int32_t e0,
                             const float* srcpos, float* bary) {
    const int32_t* e = elem + 4 * (size_t)(e0 - 1);
    f3 s0 = {srcpos[0], srcpos[1], srcpos[2]};
    for (int j = 0; j < 4; j++) {
        f3 A = at(node, e[kOut[j][0]]), B = at(node, e[kOut[j][1]]), C = at(node, e[kOut[j][2]]);
        f3 N = crs(sub(A, B), sub(A, C));
        bary[kFaceMap[j]] = -dt(sub(A, s0), N);
    }
    float s = 0.f;
    for (int j = 0; j < 4; j++) {
        if (bary[j] < 0.f) return 1;
        s += bary[j];
    }
    for (int j = 0; j < 4; j++) bary[j] /= s;
    return 0;
}

int32_t mmc_host_initelem(int32_t ne, const float* node, const int32_t* elem,
                          const float* srcpos, float* bary) {
    for (int32_t i = 0; i < ne; i++) {
        const int32_t* e = elem + 4 * (size_t)i;
        float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
        for (int j = 0; j < 4; j++) {
            const float* p = node + 3 * (size_t)(e[j] - 1);
            for (int k = 0; k < 3; k++) {
                lo[k] = fminf(lo[k], p[k]);
                hi[k] = fmaxf(hi[k], p[k]);
            }
        }
        int inbox = 1;
        for (int k = 0; k < 3; k++) inbox &= (srcpos[k] >= lo[k]) & (srcpos[k] <= hi[k]);
        if (inbox && mmc_host_barycentric(node, elem, i + 1, srcpos, bary) == 0) return i + 1;
    }
    return 0;
}

double mmc_host_reff(double n_in, double n_out) {
    const int count = 1000;
    const double step = M_PI / (2.0 * count);
    const double crit = asin(1.0 / n_in);
    double rphi = 0.0, rj = 0.0;
    for (int i = 0; i < count; i++) {
        double o = i * step, co = cos(o), so = sin(o), rf;
        if (o < crit) {
            double ct = n_in * so;
            ct = sqrt(1. - ct * ct);
            double rs = (n_in * ct - n_out * co) / (n_in * ct + n_out * co);
            double rp = (n_in * co - n_out * ct) / (n_in * co + n_out * ct);
            rf = 0.5 * rs * rs;
            rf += 0.5 * rp * rp;
        } else {
            rf = 1.f;
        }
        rphi += 2.0 * so * co * rf;
        rj += 3.0 * so * co * co * rf;
    }
    rphi *= step;
    rj *= step;
    return (rphi + rj) / (2.0 - rphi + rj);
}

void mmc_host_surface_nvol(int32_t nn, int32_t ne, const int32_t* elem, const int32_t* type,
                           const int32_t* facenb, const float* med, int32_t prop,
                           int32_t isreflect, float* nvol) {
    float* reff = (float*)calloc((size_t)prop + 1, sizeof(float));
    if (isreflect) {
        for (int32_t i = 1; i <= prop; i++) {
            for (int32_t j = 1; j < i; j++)
                if (med[4 * j + 3] == med[4 * i + 3]) {
                    reff[i] = reff[j];
                    break;
                }
            if (reff[i] == 0.f) reff[i] = (float)mmc_host_reff(med[4 * i + 3], med[3]);
        }
    }
    for (int32_t i = 0; i < ne; i++) {
        for (int j = 0; j < 4; j++) {
            if (facenb[4 * (size_t)i + j] != 0) continue;
            const int* fn = kOut[kIFaceOrder[j]];
            for (int k = 0; k < 3; k++) {
                int32_t nid = elem[4 * (size_t)i + fn[k]] - 1;
                if (nvol[nid] > 0.f && type[i] >= 0) nvol[nid] *= -(2.f / (1.0 + reff[type[i]]));
            }
        }
    }
    free(reff);
    for (int32_t i = 0; i < nn; i++)
        if (nvol[i] < 0.f) nvol[i] = -nvol[i];
}

int32_t mmc_host_number_exterior(int32_t ne, int32_t* facenb) {
    int32_t nf = 0;
    for (size_t k = 0; k < 4 * (size_t)ne; k++)
        if (facenb[k] == 0) facenb[k] = -(++nf);
    return nf;
}

int32_t mmc_host_srcelem(int32_t ne, int32_t* type, int32_t* srcelem) {
    int32_t n = 0;
    for (int32_t i = 0; i < ne; i++)
        if (type[i] == -1) {
            if (srcelem) srcelem[n] = i + 1;
            type[i] = 0;
            n++;
        }
    return n;
}

/* The reference GPU backend's nodal output under the Badouel tracer
 * (src/mmc_cl_host.c:973-1024): each element value of every gate and pattern
 * goes a quarter to each of its four nodes, added into fnode
 * ([gate][node][pattern]) from felem ([gate][element][pattern]). */
int32_t mmc_host_elem_to_nodes(int32_t ne, int32_t nn, const int32_t* elem, int32_t maxgate, int32_t srcnum,
                               const double* felem, double* fnode) {
    if (ne < 0 || nn < 1 || maxgate < 1 || srcnum < 1 || !elem || !felem || !fnode) return -1;
    for (int32_t g = 0; g < maxgate; g++)
        for (int32_t e = 0; e < ne; e++)
            for (int32_t k = 0; k < srcnum; k++) {
                const double w = felem[((size_t)g * ne + e) * srcnum + k] * 0.25;
                for (int j = 0; j < 4; j++) {
                    const int32_t v = elem[4 * (size_t)e + j];
                    if (v < 1 || v > nn) return -1;
                    fnode[((size_t)g * nn + (v - 1)) * srcnum + k] += w;
                }
            }
    return 0;
}

int32_t mmc_host_detelem(int32_t ne, int32_t* type, int32_t prop) {
    int32_t n = 0;
    for (int32_t i = 0; i < ne; i++)
        if (type[i] == -2) {
            type[i] = prop + 1;
            n++;
        }
    return n;
}

/* Gradients of the four linear basis functions of a tetrahedron: phi_i is 0
 * on the face opposite node i and 1 at node i, so grad(phi_i) is the face's
 * normal c = (b - a) x (c - a) over its projection onto (p_i - a). */
static void basis_gradients(const float* node, const int32_t* ee, double g[4][3]) {
    static const int opp[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int i = 0; i < 4; i++) {
        const float* p = node + 3 * (size_t)(ee[i] - 1);
        const float* a = node + 3 * (size_t)(ee[opp[i][0]] - 1);
        const float* b = node + 3 * (size_t)(ee[opp[i][1]] - 1);
        const float* c = node + 3 * (size_t)(ee[opp[i][2]] - 1);
        const double u[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
        const double v[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
        const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double h = n[0] * ((double)p[0] - a[0]) + n[1] * ((double)p[1] - a[1]) + n[2] * ((double)p[2] - a[2]);
        for (int k = 0; k < 3; k++) g[i][k] = (h != 0.0) ? n[k] / h : 0.0;
    }
}

int32_t mmc_host_deldotdel(const float* node, int32_t nn, const int32_t* elem, int32_t ne, const float* evol,
                           double* ddd) {
    if (!node || !elem || !evol || !ddd || nn < 1 || ne < 0) return -1;
    for (int32_t t = 0; t < ne; t++) {
        const int32_t* ee = elem + 4 * (size_t)t;
        for (int k = 0; k < 4; k++)
            if (ee[k] < 1 || ee[k] > nn) return -1;
        double g[4][3];
        basis_gradients(node, ee, g);
        int idx = 0;
        for (int i = 0; i < 4; i++)
            for (int j = i; j < 4; j++)
                ddd[10 * (size_t)t + idx++] =
                    (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]) * (double)evol[t];
    }
    return 0;
}

int32_t mmc_host_grid(const float* node, int32_t nn, float step, float nmin[3], int32_t dim[3]) {
    if (!node || nn < 1 || !(step > 0.f) || !nmin || !dim) return -1;
    for (int k = 0; k < 3; k++) {
        float lo = 1e30f, hi = -1e30f;              /* VERY_BIG, src/mmc_mesh.h:74 */
        for (int32_t i = 0; i < nn; i++) {
            const float v = node[3 * (size_t)i + k];
            lo = (v < lo) ? v : lo;
            hi = (v > hi) ? v : hi;
        }
        lo -= 1e-6f;                                 /* EPS, src/mmc_mesh.h:70 */
        hi += 1e-6f;
        const float q = (hi - lo) / step;
        if (!(q < 2.0e9f)) return -1;
        nmin[k] = lo;
        dim[k] = (int32_t)q + 1;
    }
    return 0;
}

double mmc_host_normalize_grid(const mmc_hip_sim* sim, int64_t ncells, double* field, double absorbweight,
                               double launchweight, int32_t pair) {
    (void)absorbweight;
    float Etotal = (float)launchweight;
    const int32_t S = (sim->srctype == MMC_SRC_PATTERN && sim->srcnum > 1) ? sim->srcnum : 1;
    if (pair < 0 || pair >= S || ncells < 1) return 0.0;
    const float u = sim->unitinmm > 0.f ? sim->unitinmm : 1.f;
    double k;
    if (sim->outputtype == MMC_OUT_JACOBIAN || sim->outputtype == MMC_OUT_WL || sim->outputtype == MMC_OUT_WP)
        k = (sim->outputtype == MMC_OUT_JACOBIAN) ? 1.f / (1e-4f * (float)sim->nphoton) : 1.f / Etotal;
    else if (sim->outputtype == MMC_OUT_ENERGY)
        k = 1.f / Etotal;
    else {
        k = 1.0 / (Etotal * u * u * u);
        if (sim->outputtype == MMC_OUT_FLUX) k /= sim->tstep;
    }
    const size_t n = (size_t)ncells * (size_t)sim->maxgate;
    for (size_t i = 0; i < n; i++) field[i * S + pair] *= k;
    return k;
}

double mmc_host_normalize_pair(const mmc_hip_sim* sim, int32_t nn, int32_t ne, const int32_t* elem,
                               const int32_t* type, const float* med, const float* evol,
                               const float* nvol, double* field, double absorbweight,
                               double launchweight, int32_t pair) {
    /* mesh_normalize takes the energies as float (src/mmc_mesh.c:2154) */
    float Eabsorb = (float)absorbweight, Etotal = (float)launchweight;
    if (sim->method == MMC_RT_BLBADOUELGRID) return 0.0;   /* voxel fields: mmc_host_normalize_grid */
    const int32_t datalen = sim->basisorder ? nn : ne;
    const int32_t S = (sim->srctype == MMC_SRC_PATTERN && sim->srcnum > 1) ? sim->srcnum : 1;
    if (pair < 0 || pair >= S) return 0.0;
    /* value j of gate g of this pattern slot: field[(g*datalen + j)*S + pair] */
#define F(g, j) field[((size_t)(g) * datalen + (j)) * S + pair]
    double deposit = 0.0, k;
    if (sim->outputtype == MMC_OUT_JACOBIAN || sim->outputtype == MMC_OUT_WL || sim->outputtype == MMC_OUT_WP) {
        /* replay outputs: per replayed photon for the Jacobian, per unit of
         * detected weight for the weighted path and scattering count */
        k = (sim->outputtype == MMC_OUT_JACOBIAN) ? 1.f / (1e-4f * (float)sim->nphoton) : 1.f / Etotal;
        for (int32_t g = 0; g < sim->maxgate; g++)
            for (int32_t j = 0; j < datalen; j++) F(g, j) *= k;
        return k;
    }
    if (sim->outputtype == MMC_OUT_ENERGY) {
        k = 1.f / Etotal;
        for (int32_t g = 0; g < sim->maxgate; g++)
            for (int32_t j = 0; j < datalen; j++) F(g, j) *= k;
        return k;
    }
    if (sim->basisorder) {
        for (int32_t g = 0; g < sim->maxgate; g++)
            for (int32_t j = 0; j < nn; j++)
                if (nvol[j] > 0.f) F(g, j) /= nvol[j];
        for (int32_t i = 0; i < ne; i++) {
            const int32_t* e = elem + 4 * (size_t)i;
            double sum = 0.0;
            for (int32_t g = 0; g < sim->maxgate; g++)
                for (int c = 0; c < 4; c++) sum += (float)F(g, e[c] - 1);
            deposit += sum * evol[i] * med[4 * type[i]];
        }
        k = Eabsorb / (Etotal * deposit * 0.25f);
    } else {
        for (int32_t i = 0; i < ne; i++)
            for (int32_t g = 0; g < sim->maxgate; g++) deposit += F(g, i);
        for (int32_t i = 0; i < ne; i++) {
            double em = evol[i] * med[4 * type[i]];
            for (int32_t g = 0; g < sim->maxgate; g++) F(g, i) /= em;
        }
        k = Eabsorb / (Etotal * deposit);
    }
    if (sim->outputtype == MMC_OUT_FLUX) k /= sim->tstep;
    for (int32_t g = 0; g < sim->maxgate; g++)
        for (int32_t j = 0; j < datalen; j++) F(g, j) *= k;
#undef F
    return k;
}

double mmc_host_normalize(const mmc_hip_sim* sim, int32_t nn, int32_t ne, const int32_t* elem,
                          const int32_t* type, const float* med, const float* evol,
                          const float* nvol, double* field, double absorbweight,
                          double launchweight) {
    return mmc_host_normalize_pair(sim, nn, ne, elem, type, med, evol, nvol, field, absorbweight,
                                   launchweight, 0);
}

int32_t mmc_host_save_dat(const char* fname, const double* field, int32_t datalen,
                          int32_t maxgate) {
    FILE* fp = fopen(fname, "wt");
    if (!fp) return -1;
    for (int32_t g = 0; g < maxgate; g++)
        for (int32_t j = 0; j < datalen; j++)
            if (fprintf(fp, "%d\t%e\n", j + 1, field[(size_t)g * datalen + j]) <= 0) {
                fclose(fp);
                return -2;
            }
    fclose(fp);
    return 0;
}

int32_t mmc_host_save_bin(const char* fname, const double* field, int64_t count) {
    FILE* fp = fopen(fname, "wb");
    if (!fp) return -1;
    size_t n = fwrite(field, sizeof(double), (size_t)count, fp);
    fclose(fp);
    return n == (size_t)count ? 0 : -2;
}

int32_t mmc_host_save_mch(const char* fname, const mmc_mch_header* his, const float* data,
                          int64_t count) {
    FILE* fp = fopen(fname, "wb");
    if (!fp) return -1;
    size_t ok = fwrite(his, sizeof(mmc_mch_header), 1, fp);
    size_t n = (size_t)count * his->colcount;
    if (n) ok &= (fwrite(data, sizeof(float), n, fp) == n);
    fclose(fp);
    return ok ? 0 : -2;
}
