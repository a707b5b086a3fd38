"""Adjoint Jacobians on the dual grid (the reference's -O A / D / W outputs,
src/mmc_core.cl:2228-2400, src/mmc_cl_host.c:1331-1462).

The adjoint method needs one forward field per source and one per detector,
the detector run as a source pointing into the medium; ``forward_fields``
runs them as one batch on a resident mesh and normalizes each, and
``jacobian_grid`` forms, per voxel,

    J_mua[s][d] = -unitinmm^3 * cw_s * cw_d
    J_D[s][d]   = -unitinmm * grad(cw_s) . grad(cw_d)

(cw: the field summed over gates; complex products for RF fields) on the
device through mmc_hip_adjoint_grid.  ``jacobian_mesh`` forms the mesh-mode
Jacobians from nodal fields (the linear-basis integrals over each element,
src/mmc_core.cl:2403-2649) through mmc_hip_adjoint_mesh.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import f32p, f64p, i32p, ptr
from .mesh import TetMesh
from .sim import ResidentMesh, SimConfig, normalize_field


@dataclasses.dataclass
class Jacobians:
    mua: Optional[np.ndarray]          # [s][d][z][y][x]
    mua_im: Optional[np.ndarray]
    d: Optional[np.ndarray]
    d_im: Optional[np.ndarray]


def jacobian_grid(field: np.ndarray, field_im: Optional[np.ndarray], ns: int, nd: int, unitinmm: float = 1.0,
                  mua: bool = True, dcoeff: bool = True) -> Jacobians:
    """field: [slot][gate][z][y][x] float32 (ns sources, then nd detectors)."""
    f = np.ascontiguousarray(field, np.float32)
    if f.ndim != 5 or f.shape[0] != ns + nd:
        raise ValueError("field must be [ns + nd][gate][z][y][x]")
    fi = None if field_im is None else np.ascontiguousarray(field_im, np.float32)
    if fi is not None and fi.shape != f.shape:
        raise ValueError("field_im must match field")
    nz, ny, nx = f.shape[2:]
    shp = (ns, nd, nz, ny, nx)
    mult = 2 if fi is not None else 1
    jm = np.zeros((mult,) + shp, np.float32) if mua else None
    jd = np.zeros((mult,) + shp, np.float32) if dcoeff else None
    dim = np.array([nx, ny, nz], np.int32)
    rc = _lib.lib().mmc_hip_adjoint_grid(ptr(f, C.c_float), ptr(fi, C.c_float) if fi is not None else C.cast(None, f32p),
                                         ptr(dim, C.c_int32), int(f.shape[1]), int(ns), int(nd), C.c_float(unitinmm),
                                         ptr(jm, C.c_float) if mua else C.cast(None, f32p),
                                         ptr(jd, C.c_float) if dcoeff else C.cast(None, f32p))
    _lib.check(rc, "mmc_hip_adjoint_grid")
    pick = lambda a, k: None if a is None or k >= len(a) else a[k]
    return Jacobians(pick(jm, 0), pick(jm, 1), pick(jd, 0), pick(jd, 1))


def deldotdel(mesh: TetMesh) -> np.ndarray:
    """<grad phi_i . grad phi_j> * evol per element, packed
    [00,01,02,03,11,12,13,22,23,33] (float64, mesh_deldotdel)."""
    out = np.zeros((mesh.ne, 10), np.float64)
    rc = _lib.lib().mmc_host_deldotdel(ptr(mesh.node, C.c_float), int(mesh.nn), ptr(mesh.elem, C.c_int32),
                                       int(mesh.ne), ptr(mesh.evol, C.c_float), ptr(out, C.c_double))
    if rc != 0:
        raise ValueError("element node ids outside 1..nn")
    return out


MESH_MODES = {"fem": 0, "nodal": 1, "elem": 2}


def jacobian_mesh(mesh: TetMesh, field: np.ndarray, field_im: Optional[np.ndarray], ns: int, nd: int,
                  mode: str = "fem", mua: bool = True, dcoeff: bool = True) -> Jacobians:
    """field: nodal fields [slot][gate][node] float32 (ns sources, then nd
    detectors).  mode "fem": per-element linear-basis integrals, a quarter to
    each node (the reference default); "nodal": J_mua = -nvol cw_s cw_d
    (adjointmode 1); "elem": the per-element values.  Outputs [s][d][node]
    (or [s][d][elem])."""
    if not getattr(mesh, "prepared", False):
        raise ValueError("prepare() the mesh first")
    f = np.ascontiguousarray(field, np.float32)
    if f.ndim != 3 or f.shape[0] != ns + nd or f.shape[2] != mesh.nn:
        raise ValueError("field must be [ns + nd][gate][node]")
    fi = None if field_im is None else np.ascontiguousarray(field_im, np.float32)
    if fi is not None and fi.shape != f.shape:
        raise ValueError("field_im must match field")
    m = MESH_MODES[mode]
    ddd = np.ascontiguousarray(deldotdel(mesh), np.float32) if dcoeff else None
    n = mesh.ne if m == 2 else mesh.nn
    mult = 2 if fi is not None else 1
    jm = np.zeros((mult, ns, nd, n), np.float32) if mua else None
    jd = np.zeros((mult, ns, nd, n), np.float32) if dcoeff else None
    rc = _lib.lib().mmc_hip_adjoint_mesh(ptr(f, C.c_float), ptr(fi, C.c_float), int(mesh.nn), int(mesh.ne),
                                         ptr(mesh.elem, C.c_int32), ptr(mesh.evol, C.c_float),
                                         ptr(ddd, C.c_float), ptr(mesh.nvol, C.c_float), int(f.shape[1]),
                                         int(ns), int(nd), m, ptr(jm, C.c_float), ptr(jd, C.c_float))
    _lib.check(rc, "mmc_hip_adjoint_mesh")
    pick = lambda a, k: None if a is None or k >= len(a) else a[k]
    return Jacobians(pick(jm, 0), pick(jm, 1), pick(jd, 0), pick(jd, 1))


def detector_sources(cfg: SimConfig, detpos: Sequence[Sequence[float]],
                     detdir: Sequence[Sequence[float]]) -> List[SimConfig]:
    """One simulation per detector, launched from its position along its
    direction (cfg->detdir, w = focus) as a disk of the detector's radius
    (detpos w; a pencil when 0), with the settings of cfg -- the adjoint
    sources of src/mmc_utils.c:3760-3800."""
    out = []
    for p, v in zip(detpos, detdir):
        r = float(p[3]) if len(p) > 3 else 0.0
        v4 = tuple(v[:3]) + ((float(v[3]),) if len(v) > 3 else (0.0,))
        out.append(dataclasses.replace(cfg, srctype="disk" if r > 0 else "pencil", srcpos=tuple(p[:3]),
                                       srcdir=v4, srcparam1=(r, 0.0, 0.0, 0.0), srcparam2=(0.0, 0.0, 0.0, 0.0),
                                       srcpattern=None, e0=0, issavedet=0, detpos=None))
    return out


def forward_fields(mesh: TetMesh, sources: List[SimConfig], detectors: List[SimConfig], tune=None):
    """Normalized fields of the sources then the detectors, as float32
    [slot][gate][z][y][x] for dual-grid simulations or [slot][gate][node] for
    nodal ones (and the imaginary parts for RF)."""
    cfgs = list(sources) + list(detectors)
    grid = all(c.is_grid for c in cfgs)
    if not grid and not all(not c.is_grid and c.basisorder == 1 for c in cfgs):
        raise ValueError("adjoint Jacobians need all dual-grid (method 'G') or all nodal (basisorder 1) fields")
    rm = ResidentMesh(mesh)
    try:
        res = rm.run_batch(cfgs, tune=tune)
    finally:
        rm.release()
    shape = tuple(cfgs[0].grid(mesh)[1][::-1]) if grid else (mesh.nn,)
    re, im = [], []
    for c, r in zip(cfgs, res):
        normalize_field(mesh, c, r)
        g = c.gates()[0]
        re.append(r.field.reshape((g,) + shape))
        if r.field_im is not None:
            im.append(r.field_im.reshape((g,) + shape))
    f = np.ascontiguousarray(np.stack(re), np.float32)
    fi = np.ascontiguousarray(np.stack(im), np.float32) if im else None
    return f, fi
