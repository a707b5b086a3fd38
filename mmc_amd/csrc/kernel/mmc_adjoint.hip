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
 */
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include "mmc_hip.h"

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
    if (!field || !dim || maxgate < 1 || ns < 1 || nd < 1 || (!jmua && !jd)) return -1;
    if (dim[0] < 1 || dim[1] < 1 || dim[2] < 1) return -1;
    const uint64_t ncell = (uint64_t)dim[0] * dim[1] * dim[2];
    const uint64_t nslot = (uint64_t)ns + nd;
    if ((uint64_t)maxgate * nslot >= (1ull << 32) || ncell >= (1ull << 40)) return -1;
    const uint64_t nin = ncell * maxgate * nslot, nout = ncell * ns * nd * (field_im ? 2 : 1);
    Dev dre, dim_, dmua, dd;
    hipError_t e = hipMalloc(&dre.p, sizeof(float) * nin);
    if (!e && field_im) e = hipMalloc(&dim_.p, sizeof(float) * nin);
    if (!e && jmua) e = hipMalloc(&dmua.p, sizeof(float) * nout);
    if (!e && jd) e = hipMalloc(&dd.p, sizeof(float) * nout);
    if (!e) e = hipMemcpy(dre.p, field, sizeof(float) * nin, hipMemcpyHostToDevice);
    if (!e && field_im) e = hipMemcpy(dim_.p, field_im, sizeof(float) * nin, hipMemcpyHostToDevice);
    if (e) return -100 - (int)e;
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
    return e ? -100 - (int)e : 0;
}
