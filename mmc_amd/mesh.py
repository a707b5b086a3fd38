"""Tetrahedral mesh and its preparation for transport.

``TetMesh.prepare`` runs the host-side chain of the reference's mmc_prep()
(src/mmc_host.c:136-165) through the C functions of include/mmc_hip_host.h:
orientation fix and volumes (mesh_getvolume), face neighbours
(mesh_getfacenb), the surface-node volume correction and exterior-face
numbering (tracer_prep) and the face planes (tracer_build).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import MeshC, f32p, i32p, ptr


class TetMesh:
    """Mesh arrays with tetmesh conventions (1-based ids, media row 0 = background)."""

    def __init__(self, node, elem, etype, med, srcelem=None):
        self.node = np.ascontiguousarray(node, np.float32).reshape(-1, 3)
        self.elem = np.ascontiguousarray(elem, np.int32).reshape(-1, 4).copy()
        self.etype = np.ascontiguousarray(etype, np.int32).reshape(-1)
        self.med = np.ascontiguousarray(med, np.float32).reshape(-1, 4)
        self.srcelem = None if srcelem is None else np.ascontiguousarray(srcelem, np.int32)
        self.nn, self.ne = len(self.node), len(self.elem)
        self.prop = len(self.med) - 1
        if self.elem.min() < 1 or self.elem.max() > self.nn:
            raise ValueError("element node ids outside 1..nn")
        # source elements (label -1): candidates of the area sources' launch
        # search, relabelled 0 (mmc_host_srcelem; src/mmc_mesh.c:390-427)
        if (self.etype == -1).any():
            self.etype = self.etype.copy()
            ids = np.zeros(self.ne, np.int32)
            n = _lib.lib().mmc_host_srcelem(self.ne, ptr(self.etype, _lib.C.c_int32), ptr(ids, _lib.C.c_int32))
            if self.srcelem is None:
                self.srcelem = ids[:n].copy()
        # wide-field detector elements: label -2 becomes prop+1 with a copy of
        # the background medium (mmc_host_detelem; src/mmc_mesh.c:390-427, 511-527)
        self.etype = self.etype.copy()
        self.isextdet = int(_lib.lib().mmc_host_detelem(self.ne, ptr(self.etype, _lib.C.c_int32), self.prop) > 0)
        if self.isextdet:
            self.med = np.ascontiguousarray(np.vstack([self.med, self.med[:1]]), np.float32)
        if self.etype.min() < 0 or self.etype.max() > self.prop + self.isextdet:
            raise ValueError("element labels outside 0..prop")
        self.prepared = False

    # ------------------------------------------------------------------ prep
    def prepare(self, nodal: bool = False, isreflect: int = 1, isnormalized: int = 1):
        """Orientation, volumes, neighbours, planes; nodal=True applies the
        surface-node volume correction used by nodal normalization."""
        L = _lib.lib()
        self.evol = np.zeros(self.ne, np.float32)
        self.nvol = np.zeros(self.nn, np.float32)
        L.mmc_host_volume(self.nn, ptr(self.node, _lib.C.c_float), self.ne, ptr(self.elem, _lib.C.c_int32),
                          ptr(self.etype, _lib.C.c_int32), ptr(self.evol, _lib.C.c_float),
                          ptr(self.nvol, _lib.C.c_float))
        fnb = np.zeros((self.ne, 4), np.int32)
        rc = L.mmc_host_facenb(self.nn, self.ne, ptr(self.elem, _lib.C.c_int32), ptr(fnb, _lib.C.c_int32))
        _lib.check(rc, "mmc_host_facenb")
        self.facenb_raw = fnb.copy()
        if nodal and isnormalized == 1:
            L.mmc_host_surface_nvol(self.nn, self.ne, ptr(self.elem, _lib.C.c_int32),
                                    ptr(self.etype, _lib.C.c_int32), ptr(fnb, _lib.C.c_int32),
                                    ptr(self.med, _lib.C.c_float), self.prop, int(isreflect),
                                    ptr(self.nvol, _lib.C.c_float))
        self.nf = L.mmc_host_number_exterior(self.ne, ptr(fnb, _lib.C.c_int32))
        self.facenb = fnb
        self.planes = np.zeros((self.ne, 16), np.float32)
        L.mmc_host_planes(self.ne, ptr(self.node, _lib.C.c_float), ptr(self.elem, _lib.C.c_int32),
                          ptr(self.planes, _lib.C.c_float))
        self.prepared = True
        self.nodal = nodal
        return self

    def locate(self, srcpos, e0: int = 0):
        """Initial element and barycentric weights of a point source
        (tracer_prep: e0 if it encloses the point, else mesh_initelem)."""
        L = _lib.lib()
        p = np.asarray(srcpos[:3], np.float32)
        bary = np.zeros(4, np.float32)
        if e0 > 0 and L.mmc_host_barycentric(ptr(self.node, _lib.C.c_float), ptr(self.elem, _lib.C.c_int32),
                                              int(e0), ptr(p, _lib.C.c_float), ptr(bary, _lib.C.c_float)) == 0:
            return int(e0), bary
        e = L.mmc_host_initelem(self.ne, ptr(self.node, _lib.C.c_float), ptr(self.elem, _lib.C.c_int32),
                                ptr(p, _lib.C.c_float), ptr(bary, _lib.C.c_float))
        if e <= 0:
            raise ValueError("initial element does not enclose the source")
        return int(e), bary

    def cstruct(self) -> MeshC:
        if not self.prepared:
            raise RuntimeError("mesh not prepared")
        m = MeshC()
        m.nn, m.ne, m.prop, m.nf = self.nn, self.ne, self.prop, int(self.nf)
        m.isextdet = self.isextdet
        m.node = ptr(self.node, _lib.C.c_float)
        m.elem = ptr(self.elem, _lib.C.c_int32)
        m.type = ptr(self.etype, _lib.C.c_int32)
        m.facenb = ptr(self.facenb, _lib.C.c_int32)
        m.planes = ptr(self.planes, _lib.C.c_float)
        m.med = ptr(self.med, _lib.C.c_float)
        if self.srcelem is not None and len(self.srcelem):
            m.srcelem = ptr(self.srcelem, _lib.C.c_int32)
            m.srcelemlen = len(self.srcelem)
        else:
            m.srcelem = _lib.C.cast(None, i32p)
            m.srcelemlen = 0
        m._keep = (self,)
        return m
