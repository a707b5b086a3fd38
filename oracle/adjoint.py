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


# ---------------------------------------------------------------- mesh mode
# Restates mesh_deldotdel (src/mmc_mesh.c:990-1046), mmc_adjoint_mesh_full_kernel
# and mmc_adjoint_mesh_nodal_kernel (src/mmc_core.cl:2403-2649) as used by
# src/mmc_cl_host.c:1117-1320 (nodal output, a quarter of each element value to
# each of its nodes).

PAIRS = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
DIAG = [0, 4, 7, 9]
OFF = [1, 2, 3, 5, 6, 8]


def deldotdel(node, elem, evol) -> np.ndarray:
    """<grad phi_i . grad phi_j> * evol per element, packed [00,01,02,03,11,12,13,22,23,33]
    (float64).  The gradients are the rows 1..3 of the inverse of the 4x4
    matrix [1 x y z] of the element's nodes (phi_i(x) = [1 x y z] . column i)."""
    p = np.asarray(node, np.float64)[np.asarray(elem) - 1]            # ne x 4 x 3
    a = np.concatenate([np.ones(p.shape[:2] + (1,)), p], axis=2)      # ne x 4 x 4
    g = np.linalg.inv(a)[:, 1:, :]                                    # ne x 3 x 4: column i = grad phi_i
    out = np.zeros((len(p), 10))
    k = 0
    for i in range(4):
        for j in range(i, 4):
            out[:, k] = (g[:, :, i] * g[:, :, j]).sum(1) * np.asarray(evol, np.float64)
            k += 1
    return out


def mesh_element_values(fre, fim, elem, evol, ddd, ns: int, nd: int):
    """Per element (J_mua re, im, J_D re, im), [s][d][elem], float32, in the
    operation order of the element kernel; im entries None without fim."""
    ee = np.asarray(elem, np.int64) - 1
    cr = cw(np.asarray(fre, np.float32))                              # [slot][node]
    ci = cw(np.asarray(fim, np.float32)) if fim is not None else None
    w = np.asarray(ddd, np.float32)
    vm = f32(-0.1) * np.asarray(evol, np.float32)
    ne = len(ee)
    jm = np.zeros((ns, nd, ne), np.float32)
    jd = np.zeros((ns, nd, ne), np.float32)
    jmi = np.zeros_like(jm) if ci is not None else None
    jdi = np.zeros_like(jd) if ci is not None else None
    for s in range(ns):
        sr = cr[s][ee]
        si = ci[s][ee] if ci is not None else None
        for d in range(nd):
            dr = cr[ns + d][ee]
            di = ci[ns + d][ee] if ci is not None else None
            mr, mi, kr, ki = (np.zeros(ne, np.float32) for _ in range(4))
            for i in range(4):
                pre = sr[:, i] * dr[:, i]
                pim = np.zeros(ne, np.float32)
                if ci is not None:
                    pre = pre - si[:, i] * di[:, i]
                    pim = sr[:, i] * di[:, i] + si[:, i] * dr[:, i]
                mr, mi = mr + pre, mi + pim
                kr, ki = kr + w[:, DIAG[i]] * pre, ki + w[:, DIAG[i]] * pim
            for p, (a, b) in enumerate(PAIRS):
                pre = sr[:, a] * dr[:, b] + sr[:, b] * dr[:, a]
                pim = np.zeros(ne, np.float32)
                if ci is not None:
                    pre = pre - si[:, a] * di[:, b] - si[:, b] * di[:, a]
                    pim = sr[:, a] * di[:, b] + sr[:, b] * di[:, a] + si[:, a] * dr[:, b] + si[:, b] * dr[:, a]
                mr, mi = mr + f32(0.5) * pre, mi + f32(0.5) * pim
                kr, ki = kr + w[:, OFF[p]] * pre, ki + w[:, OFF[p]] * pim
            jm[s, d], jd[s, d] = mr * vm, -kr
            if ci is not None:
                jmi[s, d], jdi[s, d] = mi * vm, -ki
    return jm, jmi, jd, jdi


def to_nodes(ev: np.ndarray, elem, nn: int) -> np.ndarray:
    """A quarter of each element's value to each of its nodes, summed in
    element order (float32): [..][elem] -> [..][node]."""
    ee = np.asarray(elem, np.int64).reshape(-1) - 1
    lead = ev.shape[:-1]
    flat = ev.reshape(-1, ev.shape[-1])
    out = np.zeros((len(flat), nn), np.float32)
    for b in range(len(flat)):
        np.add.at(out[b], ee, np.repeat(f32(0.25) * flat[b], 4))
    return out.reshape(lead + (nn,))


def mesh_nodal_approx(fre, fim, nvol, ns: int, nd: int):
    """J_mua(n) = -nvol(n) cw_s(n) cw_d(n) (re, im)."""
    cr = cw(np.asarray(fre, np.float32))
    ci = cw(np.asarray(fim, np.float32)) if fim is not None else None
    v = -np.asarray(nvol, np.float32)
    jm = np.zeros((ns, nd, cr.shape[1]), np.float32)
    jmi = np.zeros_like(jm) if ci is not None else None
    for s in range(ns):
        for d in range(nd):
            k = ns + d
            if ci is None:
                jm[s, d] = v * cr[s] * cr[k]
            else:
                jm[s, d] = v * (cr[s] * cr[k] - ci[s] * ci[k])
                jmi[s, d] = v * (cr[s] * ci[k] + ci[s] * cr[k])
    return jm, jmi
