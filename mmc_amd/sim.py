"""Simulation settings and the transport entry points (plain and batched).

``SimConfig`` carries the mcconfig fields of the photon-transport path
(src/mmc_utils.h:210-345) with the reference's defaults (mcx_initcfg,
src/mmc_utils.c:207-363) and its time-gate rule (mcx_validatecfg,
src/mmc_utils.c:3500-3530).  ``run`` is the plain drop-in (mmc_hip_run);
``ResidentMesh.run_batch`` runs many simulations in one persistent launch.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import PHOTON_DTYPE, ResultC, SimC, TuneC, f32p, f64p, i64p, ptr
from .mesh import TetMesh

METHODS = {"P": 0, "S": 3, "G": 4, "plucker": 0, "blbadouel": 3, "grid": 4}
OUTPUTS = {"x": 0, "f": 1, "e": 2, "j": 3, "l": 4, "p": 5, "flux": 0, "fluence": 1, "energy": 2,
           "jacobian": 3, "wl": 4, "wp": 5}
SOURCES = {"pencil": 0, "isotropic": 1, "cone": 2, "gaussian": 3, "planar": 4, "pattern": 5, "fourier": 6,
           "arcsine": 7, "disk": 8, "fourierx": 9, "fourierx2d": 10, "zgaussian": 11, "line": 12, "slit": 13,
           "pencilarray": 14, "pattern3d": 15, "hyperboloid": 16, "ring": 17}
# sources launched at srcpos in the given element (no launch search)
POINT_SOURCES = (0, 1, 2, 7)
# one trajectory record (MMC_TRAJ_RECLEN floats)
TRAJ_DTYPE = np.dtype([("id", "<u4"), ("p", "<f4", 3), ("w", "<f4"), ("eid", "<i4")])


@dataclasses.dataclass
class SimConfig:
    nphoton: int = 1000
    seed: int = 0x623F9A9E
    photon_offset: int = 0
    method: str = "S"              # 'P' Plucker (nodal), 'S' branch-less Badouel (element),
                                   # 'G' dual grid (branch-less Badouel, voxel output)
    basisorder: Optional[int] = None
    outputtype: str = "x"
    isreflect: int = 1
    isspecular: int = 0
    voidtime: int = 1
    isnormalized: int = 1
    srctype: str = "pencil"
    srcpos: Sequence[float] = (0.0, 0.0, 0.0)
    srcdir: Sequence[float] = (0.0, 0.0, 1.0)
    srcparam1: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    srcparam2: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    srcpattern: Optional[np.ndarray] = None   # (ny, nx), or (srcnum, ny, nx) patterns sharing photons
    tstart: float = 0.0
    tend: float = 5e-9
    tstep: float = 5e-9
    nout: float = 1.0
    minenergy: float = 1e-6
    roulettesize: float = 10.0
    issavedet: int = 0
    issaveexit: int = 0
    detpos: Optional[np.ndarray] = None    # (ndet, 4): x, y, z, radius
    maxdetphoton: int = 1000000
    e0: int = 0
    med: Optional[np.ndarray] = None       # per-simulation media override
    # replay: initial xorshift128+ state of each photon (n, 2) and, for the
    # J/L/P outputs, its detected weight and time of flight (n,)
    photonseed: Optional[np.ndarray] = None
    replayweight: Optional[np.ndarray] = None
    replaytime: Optional[np.ndarray] = None
    gridstep: float = 1.0                  # dual-grid voxel edge (cfg->steps.x), mesh units
    unitinmm: float = 1.0                  # mesh length unit (normalization of grid fields)
    omega: float = 0.0                     # RF modulation (rad/s): complex fluence, field_im
    ismomentum: int = 0                    # detected records add momentum transfer per medium (-m)
    trajectory: bool = False               # record launch/scattering/end positions ("-D M")
    maxjumpdebug: int = 10_000_000         # trajectory records kept

    def __post_init__(self):
        if self.basisorder is None or METHODS[self.method] == 4:
            self.basisorder = 1 if METHODS[self.method] == 0 else 0

    @property
    def is_rf(self) -> bool:
        """RF forward: omega > 0 and not a replay (the reference's isrfforward)."""
        return self.omega > 0 and self.photonseed is None

    @property
    def is_grid(self) -> bool:
        return METHODS[self.method] == 4

    def grid(self, mesh: TetMesh):
        """(nmin, dim) of the dual grid over this mesh (mesh_createdualmesh)."""
        nmin = np.zeros(3, np.float32)
        dim = np.zeros(3, np.int32)
        node = np.ascontiguousarray(mesh.node, np.float32)
        rc = _lib.lib().mmc_host_grid(ptr(node, C.c_float), int(mesh.nn), C.c_float(self.gridstep),
                                      ptr(nmin, C.c_float), ptr(dim, C.c_int32))
        if rc != 0:
            raise ValueError(f"bad dual-grid voxel edge {self.gridstep}")
        return nmin, dim

    def datalen(self, mesh: TetMesh) -> int:
        """Values per gate: voxels (dual grid), nodes or elements."""
        if self.is_grid:
            return int(np.prod(self.grid(mesh)[1].astype(np.int64)))
        return mesh.nn if self.basisorder else mesh.ne

    @property
    def srcnum(self) -> int:
        """Patterns sharing each photon (field slots per value)."""
        if self.srctype == "pattern" and self.srcpattern is not None and np.ndim(self.srcpattern) == 3:
            return int(np.shape(self.srcpattern)[0])
        return 1

    def gates(self):
        """maxgate and the adjusted tend, in float32 as mcx_validatecfg."""
        t0, t1, dt = np.float32(self.tstart), np.float32(self.tend), np.float32(self.tstep)
        if dt > t1 - t0:
            dt = np.float32(t1 - t0)
        maxgate = int(float(np.float32((t1 - t0) / dt)) + 0.5)
        t1 = np.float32(t0 + dt * np.float32(maxgate))
        return maxgate, float(t0), float(t1), float(dt)

    def cstruct(self, mesh: TetMesh) -> SimC:
        s = SimC()
        s.nphoton, s.seed, s.photon_offset = int(self.nphoton), int(self.seed) & (2**64 - 1), int(self.photon_offset)
        s.method = METHODS[self.method]
        s.basisorder = int(self.basisorder)
        s.outputtype = OUTPUTS[self.outputtype]
        s.isreflect, s.isspecular, s.voidtime = int(self.isreflect), int(self.isspecular), int(self.voidtime)
        s.srctype = SOURCES[self.srctype]
        s.srcnum = self.srcnum
        pos = list(self.srcpos) + [0.0] * (4 - len(self.srcpos))
        dirv = list(self.srcdir) + [0.0] * (4 - len(self.srcdir))
        if s.srctype in POINT_SOURCES:
            e0, bary = mesh.locate(pos, self.e0)
        else:
            e0 = self.e0 if self.e0 > 0 else mesh.locate(pos)[0] if _encloses(mesh, pos) else 1
            bary = np.zeros(4, np.float32)
        s.e0 = e0
        s.bary0[:] = [float(b) for b in bary]
        s.srcpos[:] = pos[:4]
        s.srcdir[:] = dirv[:4]
        s.srcparam1[:] = list(self.srcparam1)[:4]
        s.srcparam2[:] = list(self.srcparam2)[:4]
        keep = []
        if self.srcpattern is not None:
            pat = np.asarray(self.srcpattern, np.float32)
            if pat.ndim == 3 and self.srctype == "pattern":   # (srcnum, ny, nx) -> pixel-major, pattern fastest
                pat = np.transpose(pat, (1, 2, 0))
            pat = np.ascontiguousarray(pat).ravel()
            keep.append(pat)
            s.srcpattern = ptr(pat, C.c_float)
        else:
            s.srcpattern = C.cast(None, f32p)
        maxgate, t0, t1, dt = self.gates()
        s.tstart, s.tend, s.tstep, s.maxgate = t0, t1, dt, maxgate
        s.nout, s.minenergy, s.roulettesize = self.nout, self.minenergy, self.roulettesize
        if self.detpos is not None and len(self.detpos) and self.issavedet:
            det = np.ascontiguousarray(self.detpos, np.float32).reshape(-1, 4)
            keep.append(det)
            s.detpos, s.detnum, s.issavedet = ptr(det, C.c_float), len(det), 1
            s.issaveexit = int(self.issaveexit)
        else:
            s.detpos, s.detnum, s.issavedet, s.issaveexit = C.cast(None, f32p), 0, 0, 0
        s.maxdetphoton = int(self.maxdetphoton)
        if self.med is not None:
            md = np.ascontiguousarray(self.med, np.float32).reshape(-1, 4)
            if md.shape[0] != mesh.prop + 1:
                raise ValueError("media override must have prop+1 rows")
            keep.append(md)
            s.med = ptr(md, C.c_float)
        else:
            s.med = C.cast(None, f32p)
        if self.photonseed is not None:
            sd = np.ascontiguousarray(self.photonseed, np.uint64).reshape(-1, 2)
            if len(sd) < int(self.photon_offset) + int(self.nphoton):
                raise ValueError("fewer saved seeds than replayed photons")
            keep.append(sd)
            s.photonseed = sd.ctypes.data_as(C.POINTER(C.c_uint64))
        else:
            s.photonseed = C.cast(None, C.POINTER(C.c_uint64))
        for name in ("replayweight", "replaytime"):
            v = getattr(self, name)
            if v is not None:
                a = np.ascontiguousarray(v, np.float32).ravel()
                keep.append(a)
                setattr(s, name, ptr(a, C.c_float))
            else:
                setattr(s, name, C.cast(None, f32p))
        s.gridstep = float(self.gridstep)
        s.unitinmm = float(self.unitinmm)
        s.debuglevel = 8192 if self.trajectory else 0
        s.omega = float(self.omega)
        s.ismomentum = int(self.ismomentum)
        s._keep = keep
        return s

    @property
    def reclen(self):
        return None


def _encloses(mesh: TetMesh, pos) -> bool:
    try:
        mesh.locate(pos)
        return True
    except ValueError:
        return False


@dataclasses.dataclass
class SimResult:
    field: Optional[np.ndarray]
    fieldsum: float
    launchweight: float
    absorbweight: float
    escapeweight: float
    roulettegain: float
    raytet: float
    detected: int
    detdata: Optional[np.ndarray]
    detphoton: Optional[np.ndarray]
    photons: Optional[np.ndarray]
    kernel_ms: float
    total_ms: float
    normalizer: float = 1.0
    srcweight: Optional[np.ndarray] = None    # (2, srcnum): launched, absorbed energy per pattern
    traj: Optional[np.ndarray] = None         # TRAJ_DTYPE records (photon id, position, weight, element)
    field_im: Optional[np.ndarray] = None     # RF: imaginary part of the field
    ntraj: int = 0                            # records produced (more than kept if the buffer filled)


class _ResultBuf:
    def __init__(self, mesh: TetMesh, cfg: SimConfig, want_field=True, records=None):
        maxgate = cfg.gates()[0]
        datalen = cfg.datalen(mesh)
        S = cfg.srcnum
        shape = (maxgate, datalen, S) if S > 1 else (maxgate, datalen)
        self.field = np.zeros(shape, np.float64) if want_field else None
        self.field_im = np.zeros(shape, np.float64) if (want_field and cfg.is_rf) else None
        self.srcweight = np.zeros((2, S), np.float64) if S > 1 else None
        ndet = len(cfg.detpos) if (cfg.issavedet and cfg.detpos is not None) else 0
        mom = int(cfg.ismomentum > 0)
        self.reclen = (1 + (2 + mom) * mesh.prop + 6 * int(cfg.issaveexit) + 1) if ndet else 0
        nd = min(int(cfg.maxdetphoton), int(cfg.nphoton)) if ndet else 0
        self.detdata = np.zeros((max(nd, 1), max(self.reclen, 1)), np.float32) if ndet else None
        self.detphoton = np.zeros(max(nd, 1), np.int64) if ndet else None
        self.photons = None
        self.rec = records
        self.traj = np.zeros(int(cfg.maxjumpdebug), TRAJ_DTYPE) if cfg.trajectory else None
        r = ResultC()
        if self.traj is not None:
            r.traj = self.traj.ctypes.data_as(f32p)
            r.maxtraj = len(self.traj)
        r.field = ptr(self.field, C.c_double) if want_field else C.cast(None, f64p)
        r.field_im = ptr(self.field_im, C.c_double) if self.field_im is not None else C.cast(None, f64p)
        r.detdata = ptr(self.detdata, C.c_float) if ndet else C.cast(None, f32p)
        r.detphoton = ptr(self.detphoton, C.c_int64) if ndet else C.cast(None, i64p)
        r.srcweight = ptr(self.srcweight, C.c_double) if self.srcweight is not None else C.cast(None, f64p)
        if records is not None:
            first, count = records
            count = max(0, min(int(count), int(cfg.nphoton) - int(first)))
            self.photons = np.zeros(count, PHOTON_DTYPE)
            r.photons = self.photons.ctypes.data_as(C.c_void_p)
            r.recfirst, r.reccount = int(first), count
        self.c = r

    def result(self) -> SimResult:
        r = self.c
        n = int(min(r.detected, len(self.detdata))) if self.detdata is not None else 0
        return SimResult(
            field=self.field, fieldsum=r.fieldsum, launchweight=r.launchweight, absorbweight=r.absorbweight,
            escapeweight=r.escapeweight, roulettegain=r.roulettegain, raytet=r.raytet,
            detected=int(r.detected),
            detdata=self.detdata[:n] if self.detdata is not None else None,
            detphoton=self.detphoton[:n] if self.detphoton is not None else None,
            photons=self.photons, kernel_ms=r.kernel_ms, total_ms=r.total_ms, srcweight=self.srcweight,
            traj=self.traj[:min(int(r.ntraj), len(self.traj))] if self.traj is not None else None,
            ntraj=int(r.ntraj), field_im=self.field_im)


def tune_struct(block=0, blocks_per_cu=0, compensated=0, chunk=1, replicas=0) -> TuneC:
    """Launch parameters; replicas = FP32 field copies (0: automatic, up to 8
    per XCD for small fields, one per XCD when they fit the device-memory
    budget)."""
    t = TuneC()
    t.block, t.blocks_per_cu, t.chunk, t.compensated = int(block), int(blocks_per_cu), int(chunk), int(compensated)
    t.replicas = int(replicas)
    return t


def run(mesh: TetMesh, cfg: SimConfig, records=None, normalize=False, want_field=True, tune=None) -> SimResult:
    """Plain drop-in: upload, simulate, read back, free (mmc_hip_run)."""
    L = _lib.lib()
    mc = mesh.cstruct()
    sc = cfg.cstruct(mesh)
    rb = _ResultBuf(mesh, cfg, want_field, records)
    t = tune if tune is not None else tune_struct()
    _lib.check(L.mmc_hip_run(C.byref(mc), C.byref(sc), C.byref(t), C.byref(rb.c)), "mmc_hip_run")
    res = rb.result()
    if normalize and want_field and cfg.isnormalized:
        res.normalizer = normalize_field(mesh, cfg, res)
    return res


def normalize_field(mesh: TetMesh, cfg: SimConfig, res: SimResult):
    """mesh_normalize in place; with shared patterns each pattern slot is
    normalized with its own energies and the factors are returned as an array.
    An RF imaginary part is scaled by the same factor (the reference's
    imag_slot *= normalizor, src/mmc_mesh.c:2262-2274)."""
    k = _normalize_real(mesh, cfg, res)
    if res.field_im is not None:
        res.field_im *= k
    return k


def _normalize_real(mesh: TetMesh, cfg: SimConfig, res: SimResult):
    L = _lib.lib()
    sc = cfg.cstruct(mesh)
    if cfg.is_grid:
        ncells = cfg.datalen(mesh)
        if cfg.srcnum == 1:
            return L.mmc_host_normalize_grid(C.byref(sc), ncells, ptr(res.field, C.c_double), res.absorbweight,
                                             res.launchweight, 0)
        return np.array([L.mmc_host_normalize_grid(C.byref(sc), ncells, ptr(res.field, C.c_double),
                                                   res.srcweight[1, k], res.srcweight[0, k], k)
                         for k in range(cfg.srcnum)])
    args = (C.byref(sc), mesh.nn, mesh.ne, ptr(mesh.elem, C.c_int32), ptr(mesh.etype, C.c_int32),
            ptr(mesh.med, C.c_float), ptr(mesh.evol, C.c_float), ptr(mesh.nvol, C.c_float),
            ptr(res.field, C.c_double))
    if cfg.srcnum == 1:
        return L.mmc_host_normalize(*args, res.absorbweight, res.launchweight)
    return np.array([L.mmc_host_normalize_pair(*args, res.srcweight[1, k], res.srcweight[0, k], k)
                     for k in range(cfg.srcnum)])


class ResidentMesh:
    """A mesh kept in device memory across simulations (batching, serving)."""

    def __init__(self, mesh: TetMesh):
        self.mesh = mesh
        L = _lib.lib()
        h = C.c_void_p()
        mc = mesh.cstruct()
        _lib.check(L.mmc_hip_mesh_upload(C.byref(mc), C.byref(h)), "mmc_hip_mesh_upload")
        self.handle = h

    def release(self):
        if self.handle:
            _lib.lib().mmc_hip_mesh_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return int(_lib.lib().mmc_hip_mesh_bytes(self.handle))

    def output_bytes(self, cfg: SimConfig) -> int:
        sc = cfg.cstruct(self.mesh)
        return int(_lib.lib().mmc_hip_sim_output_bytes(self.handle, C.byref(sc)))

    def run_batch(self, cfgs: List[SimConfig], want_field=True, tune=None, records=None) -> List[SimResult]:
        L = _lib.lib()
        n = len(cfgs)
        sims = (SimC * n)()
        keep = []
        for i, c in enumerate(cfgs):
            s = c.cstruct(self.mesh)
            keep.append(s._keep)
            sims[i] = s
        bufs = [_ResultBuf(self.mesh, c, want_field, records if i == 0 else None) for i, c in enumerate(cfgs)]
        res = (ResultC * n)()
        for i, b in enumerate(bufs):
            res[i] = b.c
        t = tune if tune is not None else tune_struct()
        _lib.check(L.mmc_hip_run_batch(self.handle, sims, n, C.byref(t), res), "mmc_hip_run_batch")
        out = []
        for i, b in enumerate(bufs):
            b.c = res[i]
            out.append(b.result())
        return out


def photon_seeds(seed: int, first: int, n: int) -> np.ndarray:
    """Per-photon xorshift128+ states used by the device (host computation)."""
    out = np.zeros((n, 2), np.uint64)
    _lib.lib().mmc_hip_photon_seeds(int(seed) & (2**64 - 1), int(first), int(n), ptr(out, C.c_uint64))
    return out
