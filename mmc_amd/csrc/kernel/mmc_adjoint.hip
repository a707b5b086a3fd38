/*
 * mmc_adjoint.hip -- adjoint Jacobians on the dual grid: the products of
 * forward fields from the sources and from the detectors (each run as a
 * source) that the reference computes after an adjoint run
 * (mmc_adjoint_kernel / mmc_adjoint_dcoeff_kernel, src/mmc_core.cl:
 * 2228-2400, host scaling src/mmc_cl_host.c:1331-1462).
 *
 * Per voxel: cw = sum over gates of a slot's field (sequential, in float);
 *   J_mua[s][d] = -unitinmm^3 * cw_s * cw_d
 *   J_D[s][d]   = -unitinmm * grad(cw_s) . grad(cw_d)
 * with second-order finite differences (one-sided at the grid's faces) and,
 * for RF fields, complex products.  One thread per voxel; a voxel's field
 * values are read once per slot and gradient stencil point.
 *
 * On the tetrahedral mesh (nodal fields, mmc_adjoint_mesh_full_kernel and
 * mmc_adjoint_mesh_nodal_kernel, src/mmc_core.cl:2403-2649, host
 * src/mmc_cl_host.c:1117-1320) the products are the linear-basis integrals
 * over each element: J_mua(t) = -Ve/10 [sum_i s_i d_i + 1/2 sum_{i<j}
 * (s_i d_j + s_j d_i)] and J_D(t) = -sum_{i<=j} <grad phi_i . grad phi_j> Ve
 * (s_i d_j + s_j d_i for i<j); each node receives a quarter of the value of
 * every element it belongs to, or, in the nodal approximation, J_mua(n) =
 * -nvol(n) s_n d_n.  The quarters are gathered per node over a node-to-element
 * list in element order, so the sums do not depend on scheduling.
 */
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mmc_hip.h"
#include "mmc_kernel.h"

namespace {

__device__ __forceinline__ float cw_sum(const float* f, uint64_t vox, uint32_t slot, uint32_t maxgate,
                                        uint64_t ncell) {
    float s = 0.f;
    for (uint32_t t = 0; t < maxgate; t++) s += f[vox + (uint64_t)(t + slot * maxgate) * ncell];
    return s;
}

/* second-order difference along one axis: index i of n points, stride st */
__device__ __forceinline__ float fd_grad(const float* f, uint64_t vox, uint32_t slot, uint32_t maxgate,
                                         uint64_t ncell, uint32_t i, uint32_t n, uint64_t st) {
    if (n <= 1) return 0.f;
    const float f0 = cw_sum(f, vox, slot, maxgate, ncell);
    if (i == 0) {
        const float p1 = cw_sum(f, vox + st, slot, maxgate, ncell);
        if (n == 2) return p1 - f0;
        const float p2 = cw_sum(f, vox + 2 * st, slot, maxgate, ncell);
        return (-3.f * f0 + 4.f * p1 - p2) * 0.5f;
    }
    if (i == n - 1) {
        const float m1 = cw_sum(f, vox - st, slot, maxgate, ncell);
        if (n == 2) return f0 - m1;
        const float m2 = cw_sum(f, vox - 2 * st, slot, maxgate, ncell);
        return (m2 - 4.f * m1 + 3.f * f0) * 0.5f;
    }
    const float p1 = cw_sum(f, vox + st, slot, maxgate, ncell);
    const float m1 = cw_sum(f, vox - st, slot, maxgate, ncell);
    return (p1 - m1) * 0.5f;
}

__global__ void adjoint_grid(const float* fre, const float* fim, float* jmua, float* jd, uint32_t nx, uint32_t ny,
                             uint32_t nz, uint32_t maxgate, uint32_t ns, uint32_t nd, float smua, float sd) {
    const uint64_t ncell = (uint64_t)nx * ny * nz;
    const uint64_t vox = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (vox >= ncell) return;
    const uint64_t nxy = (uint64_t)nx * ny;
    const uint32_t ix = (uint32_t)(vox % nx), iy = (uint32_t)((vox / nx) % ny), iz = (uint32_t)(vox / nxy);
    const uint64_t alen = ncell * ns * nd;
    for (uint32_t s = 0; s < ns; s++) {
        float csr = 0.f, csi = 0.f, gs[3] = {0.f, 0.f, 0.f}, gsi[3] = {0.f, 0.f, 0.f};
        if (jmua) {
            csr = cw_sum(fre, vox, s, maxgate, ncell);
            if (fim) csi = cw_sum(fim, vox, s, maxgate, ncell);
        }
        if (jd) {
            gs[0] = fd_grad(fre, vox, s, maxgate, ncell, ix, nx, 1);
            gs[1] = fd_grad(fre, vox, s, maxgate, ncell, iy, ny, nx);
            gs[2] = fd_grad(fre, vox, s, maxgate, ncell, iz, nz, nxy);
            if (fim) {
                gsi[0] = fd_grad(fim, vox, s, maxgate, ncell, ix, nx, 1);
                gsi[1] = fd_grad(fim, vox, s, maxgate, ncell, iy, ny, nx);
                gsi[2] = fd_grad(fim, vox, s, maxgate, ncell, iz, nz, nxy);
            }
        }
        for (uint32_t d = 0; d < nd; d++) {
            const uint32_t slot = ns + d;
            const uint64_t out = vox + (uint64_t)(s * nd + d) * ncell;
            if (jmua) {
                const float cdr = cw_sum(fre, vox, slot, maxgate, ncell);
                float re = csr * cdr;
                if (fim) {
                    const float cdi = cw_sum(fim, vox, slot, maxgate, ncell);
                    re -= csi * cdi;
                    jmua[out + alen] = (csr * cdi + csi * cdr) * smua;
                }
                jmua[out] = re * smua;
            }
            if (jd) {
                const float gx = fd_grad(fre, vox, slot, maxgate, ncell, ix, nx, 1);
                const float gy = fd_grad(fre, vox, slot, maxgate, ncell, iy, ny, nx);
                const float gz = fd_grad(fre, vox, slot, maxgate, ncell, iz, nz, nxy);
                float re = gs[0] * gx + gs[1] * gy + gs[2] * gz;
                if (fim) {
                    const float hx = fd_grad(fim, vox, slot, maxgate, ncell, ix, nx, 1);
                    const float hy = fd_grad(fim, vox, slot, maxgate, ncell, iy, ny, nx);
                    const float hz = fd_grad(fim, vox, slot, maxgate, ncell, iz, nz, nxy);
                    re -= gsi[0] * hx + gsi[1] * hy + gsi[2] * hz;
                    jd[out + alen] =
                        (gs[0] * hx + gs[1] * hy + gs[2] * hz + gsi[0] * gx + gsi[1] * gy + gsi[2] * gz) * sd;
                }
                jd[out] = re * sd;
            }
        }
    }
}

/* packed upper triangle [00,01,02,03,11,12,13,22,23,33] of <grad phi_i . grad phi_j> Ve */
constexpr int kDiag[4] = {0, 4, 7, 9};
constexpr int kOff[6] = {1, 2, 3, 5, 6, 8};
constexpr int kPa[6] = {0, 0, 0, 1, 1, 2};
constexpr int kPb[6] = {1, 2, 3, 2, 3, 3};

/* one thread per element: [s][d][elem] values (imaginary block after the real one) */
__global__ void adjoint_mesh_elem(const float* fre, const float* fim, const int32_t* elem, const float* evol,
                                  const float* ddd, float* emua, float* ed, uint32_t ne, uint32_t nn,
                                  uint32_t maxgate, uint32_t ns, uint32_t nd) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne) return;
    uint64_t ee[4];
#pragma unroll
    for (int k = 0; k < 4; k++) ee[k] = (uint64_t)(elem[4 * (uint64_t)t + k] - 1);
    float w[10];
#pragma unroll
    for (int k = 0; k < 10; k++) w[k] = ed ? ddd[10 * (uint64_t)t + k] : 0.f;
    const float vm = emua ? -0.1f * evol[t] : 0.f;
    const bool rf = fim != nullptr;
    const uint64_t alen = (uint64_t)ne * ns * nd;
    for (uint32_t s = 0; s < ns; s++) {
        float sr[4], si[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            sr[k] = cw_sum(fre, ee[k], s, maxgate, nn);
            si[k] = rf ? cw_sum(fim, ee[k], s, maxgate, nn) : 0.f;
        }
        for (uint32_t d = 0; d < nd; d++) {
            float dr[4], di[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                dr[k] = cw_sum(fre, ee[k], ns + d, maxgate, nn);
                di[k] = rf ? cw_sum(fim, ee[k], ns + d, maxgate, nn) : 0.f;
            }
            float mr = 0.f, mi = 0.f, kr = 0.f, ki = 0.f;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float pre = sr[i] * dr[i], pim = 0.f;
                if (rf) {
                    pre -= si[i] * di[i];
                    pim = sr[i] * di[i] + si[i] * dr[i];
                }
                mr += pre;
                mi += pim;
                kr += w[kDiag[i]] * pre;
                ki += w[kDiag[i]] * pim;
            }
#pragma unroll
            for (int p = 0; p < 6; p++) {
                const int a = kPa[p], b = kPb[p];
                float pre = sr[a] * dr[b] + sr[b] * dr[a], pim = 0.f;
                if (rf) {
                    pre = pre - si[a] * di[b] - si[b] * di[a];
                    pim = sr[a] * di[b] + sr[b] * di[a] + si[a] * dr[b] + si[b] * dr[a];
                }
                mr += 0.5f * pre;
                mi += 0.5f * pim;
                kr += w[kOff[p]] * pre;
                ki += w[kOff[p]] * pim;
            }
            const uint64_t o = t + (uint64_t)(s * nd + d) * ne;
            if (emua) {
                emua[o] = mr * vm;
                if (rf) emua[o + alen] = mi * vm;
            }
            if (ed) {
                ed[o] = -kr;
                if (rf) ed[o + alen] = -ki;
            }
        }
    }
}

/* one thread per node: a quarter of each of its elements' values, in element order */
__global__ void adjoint_mesh_gather(const float* ev, float* out, const uint32_t* start, const uint32_t* list,
                                    uint32_t nn, uint32_t ne, uint32_t nblk) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const uint32_t k0 = start[n], k1 = start[n + 1];
    for (uint32_t b = 0; b < nblk; b++) {
        const float* e = ev + (uint64_t)b * ne;
        float acc = 0.f;
        for (uint32_t k = k0; k < k1; k++) acc += 0.25f * e[list[k]];
        out[n + (uint64_t)b * nn] = acc;
    }
}

/* nodal approximation of J_mua: -nvol(n) * cw_s(n) * cw_d(n) */
__global__ void adjoint_mesh_nodal(const float* fre, const float* fim, const float* nvol, float* jmua, uint32_t nn,
                                   uint32_t maxgate, uint32_t ns, uint32_t nd) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const float vol = nvol[n];
    const uint64_t alen = (uint64_t)nn * ns * nd;
    for (uint32_t s = 0; s < ns; s++) {
        const float csr = cw_sum(fre, n, s, maxgate, nn);
        const float csi = fim ? cw_sum(fim, n, s, maxgate, nn) : 0.f;
        for (uint32_t d = 0; d < nd; d++) {
            const float cdr = cw_sum(fre, n, ns + d, maxgate, nn);
            const uint64_t o = n + (uint64_t)(s * nd + d) * nn;
            if (fim) {
                const float cdi = cw_sum(fim, n, ns + d, maxgate, nn);
                jmua[o] = -vol * (csr * cdr - csi * cdi);
                jmua[o + alen] = -vol * (csr * cdi + csi * cdr);
            } else {
                jmua[o] = -vol * csr * cdr;
            }
        }
    }
}

struct Dev {
    void* p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int32_t mmc_hip_adjoint_grid(const float* field, const float* field_im, const int32_t dim[3],
                                        int32_t maxgate, int32_t ns, int32_t nd, float unitinmm, float* jmua,
                                        float* jd) {
    if (!field || !dim || maxgate < 1 || ns < 1 || nd < 1 || (!jmua && !jd))
        return mmc::report_error(-1, "adjoint: needs fields, maxgate, ns, nd >= 1 and an output");
    if (dim[0] < 1 || dim[1] < 1 || dim[2] < 1) return mmc::report_error(-1, "adjoint: empty grid");
    const uint64_t ncell = (uint64_t)dim[0] * dim[1] * dim[2];
    const uint64_t nslot = (uint64_t)ns + nd;
    if ((uint64_t)maxgate * nslot >= (1ull << 32) || ncell >= (1ull << 40))
        return mmc::report_error(-1, "adjoint: grid or slot count too large");
    const uint64_t nin = ncell * maxgate * nslot, nout = ncell * ns * nd * (field_im ? 2 : 1);
    Dev dre, dim_, dmua, dd;
    hipError_t e = hipMalloc(&dre.p, sizeof(float) * nin);
    if (!e && field_im) e = hipMalloc(&dim_.p, sizeof(float) * nin);
    if (!e && jmua) e = hipMalloc(&dmua.p, sizeof(float) * nout);
    if (!e && jd) e = hipMalloc(&dd.p, sizeof(float) * nout);
    if (!e) e = hipMemcpy(dre.p, field, sizeof(float) * nin, hipMemcpyHostToDevice);
    if (!e && field_im) e = hipMemcpy(dim_.p, field_im, sizeof(float) * nin, hipMemcpyHostToDevice);
    if (e) return mmc::report_error(-100 - (int)e, hipGetErrorString(e));
    /* scaling of src/mmc_cl_host.c:1436-1461: -Vvox for J_mua, -unitinmm for J_D */
    const float u = unitinmm > 0.f ? unitinmm : 1.f;
    const float vvox = u * u * u;
    const unsigned int block = 256;
    const unsigned int grid = (unsigned int)((ncell + block - 1) / block);
    hipLaunchKernelGGL(adjoint_grid, dim3(grid), dim3(block), 0, 0, (const float*)dre.p, (const float*)dim_.p,
                       (float*)dmua.p, (float*)dd.p, (uint32_t)dim[0], (uint32_t)dim[1], (uint32_t)dim[2],
                       (uint32_t)maxgate, (uint32_t)ns, (uint32_t)nd, -vvox, -u);
    e = hipGetLastError();
    if (!e && jmua) e = hipMemcpy(jmua, dmua.p, sizeof(float) * nout, hipMemcpyDeviceToHost);
    if (!e && jd) e = hipMemcpy(jd, dd.p, sizeof(float) * nout, hipMemcpyDeviceToHost);
    return e ? mmc::report_error(-100 - (int)e, hipGetErrorString(e)) : 0;
}

namespace {

template <typename T>
hipError_t upload(Dev& d, const T* h, uint64_t n) {
    hipError_t e = hipMalloc(&d.p, sizeof(T) * n);
    if (!e) e = hipMemcpy(d.p, h, sizeof(T) * n, hipMemcpyHostToDevice);
    return e;
}

}  // namespace

extern "C" int32_t mmc_hip_adjoint_mesh(const float* field, const float* field_im, int32_t nn, int32_t ne,
                                        const int32_t* elem, const float* evol, const float* deldotdel,
                                        const float* nvol, int32_t maxgate, int32_t ns, int32_t nd, int32_t mode,
                                        float* jmua, float* jd) {
    if (!field || !elem || nn < 1 || ne < 1 || maxgate < 1 || ns < 1 || nd < 1 || (!jmua && !jd))
        return mmc::report_error(-1, "adjoint: needs fields, a mesh, maxgate, ns, nd >= 1 and an output");
    if (mode < 0 || mode > 2) return mmc::report_error(-1, "adjoint: mode must be 0, 1 or 2");
    const bool approx = mode == 1, elemout = mode == 2;
    const bool fem_mua = jmua && !approx;
    if ((fem_mua && !evol) || (jd && !deldotdel) || (jmua && approx && !nvol))
        return mmc::report_error(-1, "adjoint: element volumes, gradient products or node volumes missing");
    const uint64_t nslot = (uint64_t)ns + nd;
    if ((uint64_t)maxgate * nslot >= (1ull << 32) || (uint64_t)ns * nd >= (1ull << 32))
        return mmc::report_error(-1, "adjoint: slot count too large");
    for (uint64_t k = 0; k < 4 * (uint64_t)ne; k++)
        if (elem[k] < 1 || elem[k] > nn) return mmc::report_error(-1, "adjoint: element node ids outside 1..nn");
    const int mult = field_im ? 2 : 1;
    const uint64_t nin = (uint64_t)nn * maxgate * nslot;
    const uint64_t nblk = (uint64_t)ns * nd * mult;
    const uint64_t nelemval = (uint64_t)ne * nblk, nout = (elemout ? (uint64_t)ne : (uint64_t)nn) * nblk;
    if (nblk >= (1ull << 32)) return mmc::report_error(-1, "adjoint: too many output blocks");

    Dev dre, dimg, delem, devol, dddd, dnvol, demua, ded, dstart, dlist, dmua, dd;
    hipError_t e = upload(dre, field, nin);
    if (!e && field_im) e = upload(dimg, field_im, nin);
    const bool fem = fem_mua || jd;
    if (!e && fem) e = upload(delem, elem, 4 * (uint64_t)ne);
    if (!e && fem_mua) e = upload(devol, evol, (uint64_t)ne);
    if (!e && jd) e = upload(dddd, deldotdel, 10 * (uint64_t)ne);
    if (!e && jmua && approx) e = upload(dnvol, nvol, (uint64_t)nn);
    if (!e && fem_mua) e = hipMalloc(&demua.p, sizeof(float) * nelemval);
    if (!e && jd) e = hipMalloc(&ded.p, sizeof(float) * nelemval);
    if (!e && fem && !elemout) {
        /* node -> element lists, elements in ascending order */
        uint32_t* start = (uint32_t*)calloc((size_t)nn + 1, sizeof(uint32_t));
        uint32_t* list = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (size_t)ne);
        if (!start || !list) {
            free(start);
            free(list);
            return mmc::report_error(-2, "adjoint: out of host memory");
        }
        for (uint64_t k = 0; k < 4 * (uint64_t)ne; k++) start[elem[k]]++;
        for (int32_t n = 0; n < nn; n++) start[n + 1] += start[n];
        for (int32_t t = 0; t < ne; t++)
            for (int k = 0; k < 4; k++) list[start[elem[4 * (uint64_t)t + k] - 1]++] = (uint32_t)t;
        for (int32_t n = nn; n > 0; n--) start[n] = start[n - 1];
        start[0] = 0;
        e = upload(dstart, start, (uint64_t)nn + 1);
        if (!e) e = upload(dlist, list, 4 * (uint64_t)ne);
        free(start);
        free(list);
    }
    if (!e && jmua && !elemout) e = hipMalloc(&dmua.p, sizeof(float) * nout);
    if (!e && jd && !elemout) e = hipMalloc(&dd.p, sizeof(float) * nout);
    if (e) return mmc::report_error(-100 - (int)e, hipGetErrorString(e));

    const unsigned int block = 256;
    const unsigned int gne = (unsigned int)(((uint64_t)ne + block - 1) / block);
    const unsigned int gnn = (unsigned int)(((uint64_t)nn + block - 1) / block);
    if (fem)
        hipLaunchKernelGGL(adjoint_mesh_elem, dim3(gne), dim3(block), 0, 0, (const float*)dre.p,
                           (const float*)dimg.p, (const int32_t*)delem.p, (const float*)devol.p,
                           (const float*)dddd.p, (float*)demua.p, (float*)ded.p, (uint32_t)ne, (uint32_t)nn,
                           (uint32_t)maxgate, (uint32_t)ns, (uint32_t)nd);
    if (jmua && approx)
        hipLaunchKernelGGL(adjoint_mesh_nodal, dim3(gnn), dim3(block), 0, 0, (const float*)dre.p,
                           (const float*)dimg.p, (const float*)dnvol.p, (float*)dmua.p, (uint32_t)nn,
                           (uint32_t)maxgate, (uint32_t)ns, (uint32_t)nd);
    if (fem_mua && !elemout)
        hipLaunchKernelGGL(adjoint_mesh_gather, dim3(gnn), dim3(block), 0, 0, (const float*)demua.p,
                           (float*)dmua.p, (const uint32_t*)dstart.p, (const uint32_t*)dlist.p, (uint32_t)nn,
                           (uint32_t)ne, (uint32_t)nblk);
    if (jd && !elemout)
        hipLaunchKernelGGL(adjoint_mesh_gather, dim3(gnn), dim3(block), 0, 0, (const float*)ded.p, (float*)dd.p,
                           (const uint32_t*)dstart.p, (const uint32_t*)dlist.p, (uint32_t)nn, (uint32_t)ne,
                           (uint32_t)nblk);
    e = hipGetLastError();
    const void* smua = elemout ? demua.p : dmua.p;
    const void* sd = elemout ? ded.p : dd.p;
    if (!e && jmua) e = hipMemcpy(jmua, smua, sizeof(float) * nout, hipMemcpyDeviceToHost);
    if (!e && jd) e = hipMemcpy(jd, sd, sizeof(float) * nout, hipMemcpyDeviceToHost);
    return e ? mmc::report_error(-100 - (int)e, hipGetErrorString(e)) : 0;
}
