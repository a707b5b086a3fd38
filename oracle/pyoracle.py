"""ctypes driver of the CPU oracle (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  It prepares a mesh with the oracle's own
restatement of mmc_prep (volumes, face neighbours, planes, Plucker
coordinates, surface-node correction, exterior numbering) and runs the
oracle's restatement of mmc_run_mp.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATHS = {"glibc": os.path.join(HERE, "_build", "liborc.so"),
            "portable": os.path.join(HERE, "_build", "liborc_pm.so")}

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u64p = C.POINTER(C.c_uint64)


class OMesh(C.Structure):
    _fields_ = [("nn", C.c_int), ("ne", C.c_int), ("prop", C.c_int), ("nf", C.c_int),
                ("node", f32p), ("elem", i32p), ("type", i32p), ("facenb", i32p), ("med", f32p),
                ("planes", f32p), ("pd", f32p), ("pm", f32p), ("evol", f32p), ("nvol", f32p),
                ("srcelem", i32p), ("srcelemlen", C.c_int), ("isextdet", C.c_int)]


class OCfg(C.Structure):
    _fields_ = [("nphoton", C.c_int64), ("method", C.c_int), ("basisorder", C.c_int), ("outputtype", C.c_int),
                ("isreflect", C.c_int), ("isspecular", C.c_int), ("voidtime", C.c_int), ("issavedet", C.c_int),
                ("issaveexit", C.c_int), ("ismomentum", C.c_int), ("srctype", C.c_int), ("srcnum", C.c_int),
                ("e0", C.c_int), ("bary0", C.c_float * 4), ("srcpos", C.c_float * 4), ("srcdir", C.c_float * 4),
                ("srcparam1", C.c_float * 4), ("srcparam2", C.c_float * 4), ("srcpattern", f32p),
                ("tstart", C.c_float), ("tend", C.c_float), ("tstep", C.c_float), ("maxgate", C.c_int),
                ("nout", C.c_float), ("minenergy", C.c_float), ("roulettesize", C.c_float),
                ("detnum", C.c_int), ("detpos", f32p), ("seed", C.c_uint32), ("photonseed", u64p),
                ("replayweight", f32p), ("replaytime", f32p),
                ("nthreads", C.c_int), ("recfirst", C.c_int64), ("reccount", C.c_int64),
                ("gridstep", C.c_float), ("unitinmm", C.c_float), ("gridmin", C.c_float * 3),
                ("griddim", C.c_int * 3), ("omega", C.c_float)]


class OOut(C.Structure):
    _fields_ = [("weight", f64p), ("launchweight", C.c_double), ("absorbweight", C.c_double),
                ("raytet", C.c_double), ("detdata", f32p), ("detphoton", i64p), ("maxdet", C.c_int),
                ("detected", C.c_int), ("reclen", C.c_int), ("rec", C.c_void_p), ("runtime_ms", C.c_double),
                ("error", C.c_int), ("weight_im", f64p), ("ties", i32p), ("srcweight", f64p)]


PHOTON_DTYPE = np.dtype([
    ("p", "<f4", 3), ("w", "<f4"), ("v", "<f4", 3), ("t", "<f4"),
    ("eid", "<i4"), ("reason", "<i4"), ("nraytet", "<i4"), ("detid", "<i4"),
    ("eabs", "<f4"), ("rgain", "<f4"), ("w0", "<f4"), ("nscat", "<f4"),
])

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(math: str = "glibc"):
    """The oracle library; math='glibc' (the reference's libm) or 'portable'
    (the deterministic FP32 libm shared with the GPU kernel)."""
    if math not in _libs:
        path = LIBPATHS[math]
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_rand_next.restype = C.c_float
        L.orc_rand_next.argtypes = [u64p]
        L.orc_rand_scatlen.restype = C.c_float
        L.orc_rand_scatlen.argtypes = [u64p]
        L.orc_getvolume.argtypes = [C.c_int, f32p, C.c_int, i32p, i32p, f32p, f32p]
        L.orc_getfacenb.argtypes = [C.c_int, i32p, i32p]
        L.orc_tracer_build.argtypes = [C.c_int, f32p, i32p, f32p, f32p, f32p]
        L.orc_initelem.argtypes = [C.POINTER(OMesh), f32p, f32p]
        L.orc_barycentric.argtypes = [C.POINTER(OMesh), C.c_int, f32p, f32p]
        L.orc_getreff.restype = C.c_double
        L.orc_getreff.argtypes = [C.c_double, C.c_double]
        L.orc_surface_nvol.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, f32p, C.c_int, C.c_int, f32p]
        L.orc_number_exterior.argtypes = [C.c_int, i32p]
        L.orc_run.argtypes = [C.POINTER(OCfg), C.POINTER(OMesh), C.POINTER(OOut)]
        L.orc_normalize.restype = C.c_double
        L.orc_normalize.argtypes = [C.POINTER(OCfg), C.POINTER(OMesh), f64p, C.c_float, C.c_float]
        L.orc_scatter.restype = C.c_float
        L.orc_scatter.argtypes = [C.c_float, f32p, u64p]
        L.orc_reflect.restype = C.c_float
        L.orc_reflect.argtypes = [C.POINTER(OMesh), C.c_int, C.c_float, C.c_int, C.c_int, f32p, C.c_int,
                                  i32p, C.c_int, u64p]
        L.orc_probe_raytet.argtypes = [C.POINTER(OMesh), C.c_int, C.c_int, f32p, f32p, C.c_float, f32p]
        L.orc_rand_seed.argtypes = [C.POINTER(C.c_uint32), u64p]
        _libs[math] = L
    return _libs[math]


def p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class OracleMesh:
    """A mesh prepared by the oracle's own restatement of mmc_prep."""

    def __init__(self, node, elem, etype, med, nodal=False, isreflect=1, isnormalized=1, srcelem=None):
        L = lib()
        self.node = np.ascontiguousarray(node, np.float32).reshape(-1, 3)
        self.elem = np.ascontiguousarray(elem, np.int32).reshape(-1, 4).copy()
        self.etype = np.ascontiguousarray(etype, np.int32).copy()
        self.med = np.ascontiguousarray(med, np.float32).reshape(-1, 4)
        self.nn, self.ne, self.prop = len(self.node), len(self.elem), len(self.med) - 1
        # wide-field detector elements: label -2 -> prop+1, medium = the
        # background's (mesh_srcdetelem / mesh_loadmedia, src/mmc_mesh.c:390-427, 511-527)
        self.isextdet = int(np.any(self.etype == -2))
        if self.isextdet:
            self.etype[self.etype == -2] = self.prop + 1
            self.med = np.ascontiguousarray(np.vstack([self.med, self.med[:1]]), np.float32)
        self.evol = np.zeros(self.ne, np.float32)
        self.nvol = np.zeros(self.nn, np.float32)
        L.orc_getvolume(self.nn, p(self.node, C.c_float), self.ne, p(self.elem, C.c_int32),
                        p(self.etype, C.c_int32), p(self.evol, C.c_float), p(self.nvol, C.c_float))
        self.facenb = np.zeros((self.ne, 4), np.int32)
        L.orc_getfacenb(self.ne, p(self.elem, C.c_int32), p(self.facenb, C.c_int32))
        self.facenb_raw = self.facenb.copy()
        if nodal and isnormalized == 1:
            L.orc_surface_nvol(self.nn, self.ne, p(self.elem, C.c_int32), p(self.etype, C.c_int32),
                               p(self.facenb, C.c_int32), p(self.med, C.c_float), self.prop, isreflect,
                               p(self.nvol, C.c_float))
        self.nf = L.orc_number_exterior(self.ne, p(self.facenb, C.c_int32))
        self.planes = np.zeros((self.ne, 16), np.float32)
        self.pd = np.zeros((self.ne, 18), np.float32)
        self.pm = np.zeros((self.ne, 18), np.float32)
        L.orc_tracer_build(self.ne, p(self.node, C.c_float), p(self.elem, C.c_int32), p(self.planes, C.c_float),
                           p(self.pd, C.c_float), p(self.pm, C.c_float))
        self.srcelem = None if srcelem is None else np.ascontiguousarray(srcelem, np.int32)
        m = OMesh()
        m.nn, m.ne, m.prop, m.nf = self.nn, self.ne, self.prop, self.nf
        m.node, m.elem, m.type = p(self.node, C.c_float), p(self.elem, C.c_int32), p(self.etype, C.c_int32)
        m.facenb, m.med = p(self.facenb, C.c_int32), p(self.med, C.c_float)
        m.planes, m.pd, m.pm = p(self.planes, C.c_float), p(self.pd, C.c_float), p(self.pm, C.c_float)
        m.evol, m.nvol = p(self.evol, C.c_float), p(self.nvol, C.c_float)
        m.isextdet = self.isextdet
        if self.srcelem is not None and len(self.srcelem):
            m.srcelem, m.srcelemlen = p(self.srcelem, C.c_int32), len(self.srcelem)
        else:
            m.srcelem, m.srcelemlen = C.cast(None, i32p), 0
        self.c = m

    def initelem(self, srcpos):
        bary = np.zeros(4, np.float32)
        pos = np.asarray(srcpos[:3], np.float32)
        e = lib().orc_initelem(C.byref(self.c), p(pos, C.c_float), p(bary, C.c_float))
        return e, bary


SOURCES = {"pencil": 0, "isotropic": 1, "cone": 2, "gaussian": 3, "planar": 4, "pattern": 5, "fourier": 6,
           "arcsine": 7, "disk": 8, "fourierx": 9, "fourierx2d": 10, "zgaussian": 11, "line": 12, "slit": 13,
           "pencilarray": 14, "pattern3d": 15, "hyperboloid": 16, "ring": 17}     # enum TSrcType


def make_cfg(mesh: OracleMesh, sim, photonseed=None, nthreads=0, records=None):
    """Oracle settings from an mmc_amd.SimConfig-like object (same field names)."""
    maxgate, t0, t1, dt = sim.gates()
    c = OCfg()
    c.nphoton = int(sim.nphoton)
    c.method = {"P": 0, "S": 3, "G": 4}[sim.method]
    c.basisorder = 0 if sim.method == "G" else int(sim.basisorder)
    c.unitinmm = float(getattr(sim, "unitinmm", 1.0) or 1.0)
    c.omega = float(getattr(sim, "omega", 0.0) or 0.0)
    if sim.method == "G":
        c.gridstep = float(sim.gridstep)
        c.gridmin[:], c.griddim[:] = grid(mesh, c.gridstep)
    c.outputtype = {"x": 0, "f": 1, "e": 2, "j": 3, "l": 4, "p": 5}[sim.outputtype]
    c.isreflect, c.isspecular, c.voidtime = int(sim.isreflect), int(sim.isspecular), int(sim.voidtime)
    # wide-field detector elements replace the point detectors (src/mmc_mesh.c:402-406)
    extdet = bool(getattr(mesh, "isextdet", 0))
    c.issavedet = int(bool(sim.issavedet and (extdet or (sim.detpos is not None and len(sim.detpos)))))
    c.issaveexit = int(sim.issaveexit) if c.issavedet else 0
    c.ismomentum = int(getattr(sim, "ismomentum", 0) > 0)
    c.srctype = SOURCES[sim.srctype]
    c.srcnum = int(sim.srcnum)
    pos = list(sim.srcpos) + [0.0] * (4 - len(sim.srcpos))
    dirv = list(sim.srcdir) + [0.0] * (4 - len(sim.srcdir))
    c.srcpos[:], c.srcdir[:] = pos[:4], dirv[:4]
    c.srcparam1[:], c.srcparam2[:] = list(sim.srcparam1)[:4], list(sim.srcparam2)[:4]
    keep = []
    if c.srctype in (0, 1, 2, 7):       # launched at srcpos in the given element
        e0 = 0
        bary = np.zeros(4, np.float32)
        if sim.e0 > 0:
            b = np.zeros(4, np.float32)
            pp = np.asarray(pos[:3], np.float32)
            if lib().orc_barycentric(C.byref(mesh.c), int(sim.e0), p(pp, C.c_float), p(b, C.c_float)) == 0:
                e0, bary = int(sim.e0), b
        if e0 == 0:
            e0, bary = mesh.initelem(pos)
        c.e0 = e0
        c.bary0[:] = [float(x) for x in bary]
    else:
        c.e0 = int(sim.e0) if sim.e0 > 0 else max(int(mesh.initelem(pos)[0]), 1)
    if sim.srcpattern is not None:
        pat = np.asarray(sim.srcpattern, np.float32)
        if pat.ndim == 3 and sim.srctype == "pattern":     # (srcnum, ny, nx) -> pixel-major, pattern fastest
            pat = np.transpose(pat, (1, 2, 0))
        pat = np.ascontiguousarray(pat).ravel()
        keep.append(pat)
        c.srcpattern = p(pat, C.c_float)
    else:
        c.srcpattern = C.cast(None, f32p)
    c.tstart, c.tend, c.tstep, c.maxgate = t0, t1, dt, maxgate
    c.nout, c.minenergy, c.roulettesize = sim.nout, sim.minenergy, sim.roulettesize
    if c.issavedet and not extdet:
        det = np.ascontiguousarray(sim.detpos, np.float32).reshape(-1, 4)
        keep.append(det)
        c.detnum, c.detpos = len(det), p(det, C.c_float)
    else:
        c.detnum, c.detpos = 0, C.cast(None, f32p)
    c.seed = int(sim.seed) & 0xFFFFFFFF
    if photonseed is not None:
        ps = np.ascontiguousarray(photonseed, np.uint64).reshape(-1)
        keep.append(ps)
        c.photonseed = p(ps, C.c_uint64)
    else:
        c.photonseed = C.cast(None, u64p)
    for name in ("replayweight", "replaytime"):
        v = getattr(sim, name, None)
        if v is not None:
            a = np.ascontiguousarray(v, np.float32).ravel()
            keep.append(a)
            setattr(c, name, p(a, C.c_float))
        else:
            setattr(c, name, C.cast(None, f32p))
    c.nthreads = int(nthreads)
    c.recfirst, c.reccount = (int(records[0]), int(records[1])) if records else (0, 0)
    c._keep = keep
    return c


def grid(mesh: OracleMesh, step: float):
    """Grid origin and dimensions of the dual-grid output (orc_grid)."""
    nmin = np.zeros(3, np.float32)
    dim = np.zeros(3, np.int32)
    lib().orc_grid(int(mesh.nn), p(mesh.node, C.c_float), C.c_float(step), p(nmin, C.c_float),
                   p(dim, C.c_int32))
    return [float(x) for x in nmin], [int(x) for x in dim]


class OracleResult:
    pass


def run(mesh: OracleMesh, sim, photonseed=None, nthreads=0, records=None, maxdet=0, normalize=False,
        math="glibc"):
    L = lib(math)
    cfg = make_cfg(mesh, sim, photonseed, nthreads, records)
    if cfg.method == 4:
        datalen = int(np.prod(list(cfg.griddim)))
    else:
        datalen = mesh.nn if cfg.basisorder else mesh.ne
    S = int(cfg.srcnum)
    field = np.zeros((cfg.maxgate, datalen, S) if S > 1 else (cfg.maxgate, datalen), np.float64)
    srcweight = np.zeros((2, S), np.float64) if S > 1 else None
    rf = cfg.omega > 0 and getattr(sim, "replayweight", None) is None   # RF is off for replays
    field_im = np.zeros_like(field) if rf else None
    out = OOut()
    out.weight = p(field, C.c_double)
    out.weight_im = p(field_im, C.c_double) if rf else C.cast(None, f64p)
    out.srcweight = p(srcweight, C.c_double) if srcweight is not None else C.cast(None, f64p)
    reclen = (2 + (cfg.ismomentum > 0)) * mesh.prop + 6 * cfg.issaveexit + 2
    det = np.zeros((max(maxdet, 1), reclen), np.float32)
    detph = np.zeros(max(maxdet, 1), np.int64)
    out.detdata = p(det, C.c_float) if maxdet else C.cast(None, f32p)
    out.detphoton = p(detph, C.c_int64) if maxdet else C.cast(None, i64p)
    out.maxdet = int(maxdet)
    recs = None
    ties = None
    if records:
        recs = np.zeros(int(records[1]), PHOTON_DTYPE)
        out.rec = recs.ctypes.data_as(C.c_void_p)
        ties = np.zeros((int(records[1]), 2), np.int32)
        out.ties = p(ties, C.c_int32)
    else:
        out.ties = C.cast(None, i32p)
    rc = L.orc_run(C.byref(cfg), C.byref(mesh.c), C.byref(out))
    if rc != 0:
        raise RuntimeError("oracle run failed (a launch found no enclosing element)")
    r = OracleResult()
    r.field, r.launchweight, r.absorbweight, r.raytet = field, out.launchweight, out.absorbweight, out.raytet
    n = min(out.detected, maxdet)
    r.detected, r.detdata, r.detphoton = out.detected, det[:n], detph[:n]
    r.photons, r.runtime_ms, r.reclen = recs, out.runtime_ms, out.reclen
    # per recorded photon: ray-tet tests decided by a tie rule, fixphoton nudges
    r.ties = ties
    r.field_im = field_im
    r.srcweight = srcweight
    r.normalizer = None
    if normalize:
        r.normalizer = L.orc_normalize(C.byref(cfg), C.byref(mesh.c), p(field, C.c_double),
                                       C.c_float(r.absorbweight), C.c_float(r.launchweight))
        if field_im is not None:
            field_im *= r.normalizer       # imag_slot *= normalizor (src/mmc_mesh.c:2262-2274)
    return r
