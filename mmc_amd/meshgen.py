"""Synthetic tetrahedral meshes for the benchmark workloads.

Lattice meshes split every cube cell into 6 tetrahedra sharing the cell's
000->111 diagonal (the Kuhn/Freudenthal split that the reference's
``matlab/genT6mesh.m`` builds).  Media labels are assigned by element
centroid.  Arrays follow tetmesh conventions: 1-based node ids in ``elem``,
``type`` 0..prop, media rows ``[mua, mus, g, n]`` with row 0 the background.
"""
from __future__ import annotations

import itertools

import numpy as np


def t6_lattice(nx: int, ny: int, nz: int, h: float = 1.0, origin=(0.0, 0.0, 0.0)):
    """Nodes and elements of an nx*ny*nz cell lattice, 6 tets per cell.

    Elements are ordered cell-major (the 6 tets of a cell are adjacent) with
    cells in x-fastest order.  Returns (node[nn,3] float32, elem[ne,4] int32).
    """
    nxn, nyn, nzn = nx + 1, ny + 1, nz + 1
    xs = origin[0] + h * np.arange(nxn, dtype=np.float64)
    ys = origin[1] + h * np.arange(nyn, dtype=np.float64)
    zs = origin[2] + h * np.arange(nzn, dtype=np.float64)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    node = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float32)

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = (i + nxn * (j + nyn * k)).ravel().astype(np.int64)
    step = np.array([1, nxn, nxn * nyn], dtype=np.int64)
    tets = []
    for perm in itertools.permutations(range(3)):
        v0 = np.zeros_like(base)
        v1 = v0 + step[perm[0]]
        v2 = v1 + step[perm[1]]
        v3 = v2 + step[perm[2]]
        tets.append(np.stack([base + v0, base + v1, base + v2, base + v3], axis=1))
    elem = (np.stack(tets, axis=1).reshape(-1, 4) + 1).astype(np.int32)
    # orient every tetrahedron positively (odd permutations come out negative)
    neg = signed_volumes(node, elem) < 0
    elem[neg, 2], elem[neg, 3] = elem[neg, 3].copy(), elem[neg, 2].copy()
    return node, elem


def centroids(node: np.ndarray, elem: np.ndarray) -> np.ndarray:
    return node[elem - 1].astype(np.float64).mean(axis=1)


def jitter_interior(node: np.ndarray, amp: float, seed: int, lo, hi) -> np.ndarray:
    """Perturb interior nodes by U(-amp, amp) per axis (boundary nodes fixed)."""
    rng = np.random.default_rng(seed)
    out = node.astype(np.float64).copy()
    interior = np.all((node > np.asarray(lo) + 1e-6) & (node < np.asarray(hi) - 1e-6), axis=1)
    out[interior] += rng.uniform(-amp, amp, size=(int(interior.sum()), 3))
    return out.astype(np.float32)


def signed_volumes(node: np.ndarray, elem: np.ndarray) -> np.ndarray:
    p = node[elem - 1].astype(np.float64)
    return np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0]) / 6.0


# --------------------------------------------------------------------------
# the workloads
# --------------------------------------------------------------------------

def cube60(h: float = 2.0):
    """Homogeneous 60 mm cube, h mm cells (h=2: 162,000 tets, 29,791 nodes).
    Single tissue mua=0.005/mm, mus=1.0/mm, g=0.01, n=1.37."""
    n = int(round(60 / h))
    node, elem = t6_lattice(n, n, n, h)
    etype = np.ones(len(elem), np.int32)
    med = np.array([[0, 0, 1, 1], [0.005, 1.0, 0.01, 1.37]], np.float32)
    return node, elem, etype, med


def slab_layers(h: float = 1.0, nx: int = 100, ny: int = 100, nz: int = 50):
    """100x100x50 mm box, four z-layers of 3/7/2/38 mm and a 10 mm-radius
    sphere (mua=0.05) centred at depth 20 mm; labels 1-4 layers, 5 sphere."""
    node, elem = t6_lattice(int(round(nx / h)), int(round(ny / h)), int(round(nz / h)), h)
    c = centroids(node, elem)
    z = c[:, 2]
    etype = np.where(z < 3, 1, np.where(z < 10, 2, np.where(z < 12, 3, 4))).astype(np.int32)
    r2 = (c[:, 0] - nx / 2) ** 2 + (c[:, 1] - ny / 2) ** 2 + (z - 20.0) ** 2
    etype[r2 < 10.0 ** 2] = 5
    med = np.array([[0, 0, 1, 1],
                    [0.019, 7.8, 0.89, 1.37],
                    [0.019, 7.8, 0.89, 1.37],
                    [0.004, 0.009, 0.89, 1.37],
                    [0.02, 9.0, 0.89, 1.37],
                    [0.05, 9.0, 0.89, 1.37]], np.float32)
    return node, elem, etype, med


def jittered_box(nx=120, ny=120, nz=60, h=1.0, amp=0.2, seed=20241, axes=None, nlabel=3):
    """Jittered lattice (interior nodes + U(-amp, amp)), positive volumes
    enforced by re-drawing offending nodes, labels from nested ellipsoids."""
    node, elem = t6_lattice(nx, ny, nz, h)
    lo, hi = (0, 0, 0), (nx * h, ny * h, nz * h)
    base = node.copy()
    node = jitter_interior(base, amp, seed, lo, hi)
    rng = np.random.default_rng(seed + 1)
    for _ in range(50):
        bad = signed_volumes(node, elem) <= 0
        if not bad.any():
            break
        ids = np.unique(elem[bad] - 1)
        interior = np.all((base[ids] > np.asarray(lo) + 1e-6) & (base[ids] < np.asarray(hi) - 1e-6), axis=1)
        ids = ids[interior]
        node[ids] = (base[ids] + rng.uniform(-amp, amp, size=(len(ids), 3)) * 0.5).astype(np.float32)
    else:
        raise RuntimeError("could not enforce positive volumes")
    c = centroids(node, elem)
    ctr = np.array([nx * h / 2, ny * h / 2, nz * h / 2])
    if axes is None:
        s = np.array([nx * h, ny * h, nz * h]) / 2
        axes = [s * f for f in np.linspace(0.9, 0.45, nlabel - 1)]
    etype = np.ones(len(elem), np.int32)
    for lab, ax in enumerate(axes, start=2):
        inside = (((c - ctr) / np.asarray(ax)) ** 2).sum(axis=1) < 1.0
        etype[inside] = lab
    prop = int(etype.max())
    med = np.zeros((prop + 1, 4), np.float32)
    med[0] = [0, 0, 1, 1]
    for lab in range(1, prop + 1):
        med[lab] = [0.01 * lab, 1.0 + 0.5 * lab, 0.9, 1.37]
    return node, elem, etype, med


def shells_box(nx=200, ny=200, nz=100, h=1.0, amp=0.2, seed=31,
               outer=(90.0, 80.0, 40.0), inner=(60.0, 50.0, 25.0), nshell=4):
    """Jittered box with nested ellipsoidal shells: label 1 outside the
    largest ellipsoid, labels 2..nshell+1 inside successively smaller ones
    (semi-axes linear from `outer` to `inner`), centred in the box."""
    axes = [np.asarray(outer) + (np.asarray(inner) - np.asarray(outer)) * t
            for t in np.linspace(0.0, 1.0, nshell)]
    node, elem, etype, _ = jittered_box(nx, ny, nz, h, amp=amp, seed=seed, axes=axes, nlabel=nshell + 1)
    # properties of the layered slab, and mua = 0.1/mm for the innermost shell
    med = np.array([[0, 0, 1, 1],
                    [0.019, 7.8, 0.89, 1.37],
                    [0.019, 7.8, 0.89, 1.37],
                    [0.004, 0.009, 0.89, 1.37],
                    [0.02, 9.0, 0.89, 1.37],
                    [0.1, 9.0, 0.89, 1.37]], np.float32)
    return node, elem, etype, med
