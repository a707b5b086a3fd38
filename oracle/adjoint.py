"""CPU restatement of the grid adjoint Jacobians (numpy, float32).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
mmc_hip_adjoint_grid; the product never imports it.

Restates mmc_cw_sum, mmc_fd_grad, mmc_adjoint_kernel and
mmc_adjoint_dcoeff_kernel (src/mmc_core.cl:2228-2400) and the host scaling
of src/mmc_cl_host.c:1436-1461, operation by operation in float32 (the gate
sums run sequentially, as the kernels' loops do).

Input field: [slot][gate][z][y][x] (slots: ns sources, then nd detectors);
output: [s][d][z][y][x] (real; imaginary parts separately).
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def cw(field: np.ndarray) -> np.ndarray:
    """Sum over gates, sequential in float32: [slot][gate][...] -> [slot][...]."""
    s = np.zeros(field.shape[:1] + field.shape[2:], np.float32)
    for t in range(field.shape[1]):
        s = (s + field[:, t]).astype(np.float32)
    return s


def fd_grad(c: np.ndarray, axis: int) -> np.ndarray:
    """Second-order difference of c ([slot][z][y][x]) along a spatial axis
    (0: x, 1: y, 2: z), one-sided at the faces (mmc_fd_grad)."""
    ax = 3 - axis
    n = c.shape[ax]
    g = np.zeros_like(c)
    if n <= 1:
        return g
    take = lambda i: np.take(c, i, axis=ax)
    put = lambda i, v: g.__setitem__((slice(None),) * ax + (i,), v)
    for i in range(n):
        f0 = take(i)
        if i == 0:
            p1 = take(1)
            v = p1 - f0 if n == 2 else (f32(-3) * f0 + f32(4) * p1 - take(2)) * f32(0.5)
        elif i == n - 1:
            m1 = take(n - 2)
            v = f0 - m1 if n == 2 else (take(n - 3) - f32(4) * m1 + f32(3) * f0) * f32(0.5)
        else:
            v = (take(i + 1) - take(i - 1)) * f32(0.5)
        put(i, v.astype(np.float32))
    return g


def jacobians(fre, fim, ns: int, nd: int, unitinmm: float = 1.0):
    """(J_mua re, im, J_D re, im); im entries are None without fim."""
    u = f32(unitinmm)
    smua, sd = -(u * u * u), -u
    cr = cw(np.asarray(fre, np.float32))
    ci = cw(np.asarray(fim, np.float32)) if fim is not None else None
    gr = [fd_grad(cr, a) for a in range(3)]
    gi = [fd_grad(ci, a) for a in range(3)] if ci is not None else None
    shp = (ns, nd) + cr.shape[1:]
    jm, jd = np.zeros(shp, np.float32), np.zeros(shp, np.float32)
    jmi = np.zeros(shp, np.float32) if ci is not None else None
    jdi = np.zeros(shp, np.float32) if ci is not None else None
    for s in range(ns):
        for d in range(nd):
            k = ns + d
            re = cr[s] * cr[k]
            if ci is not None:
                re = re - ci[s] * ci[k]
                jmi[s, d] = (cr[s] * ci[k] + ci[s] * cr[k]) * smua
            jm[s, d] = re * smua
            dre = gr[0][s] * gr[0][k] + gr[1][s] * gr[1][k] + gr[2][s] * gr[2][k]
            if gi is not None:
                dre = dre - (gi[0][s] * gi[0][k] + gi[1][s] * gi[1][k] + gi[2][s] * gi[2][k])
                jdi[s, d] = (gr[0][s] * gi[0][k] + gr[1][s] * gi[1][k] + gr[2][s] * gi[2][k]
                             + gi[0][s] * gr[0][k] + gi[1][s] * gr[1][k] + gi[2][s] * gr[2][k]) * sd
            jd[s, d] = dre * sd
    return jm, jmi, jd, jdi
