"""Detected-photon files and seed replay.

``save_mch``/``load_mch`` read and write the reference's ``.mch`` layout: the
64-byte ``history`` header (src/mmc_utils.h:168-184), ``savedphoton`` records
of ``colcount`` floats (``[detid, nscat[prop], ppath[prop], (p, v), w0]``)
and, when seeds are saved (``-q 1``), ``savedphoton`` xorshift128+ states of
``seedbyte = 16`` bytes (src/mmc_utils.c:4309-4342).  A photon's initial
state is a function of (seed, photon id) here, so the seeds of detected
photons come from their ids.

``replay_config`` restates mesh_loadseedfile (src/mmc_mesh.c:815-898): it
keeps the records of detector ``replaydet`` (0: all), and for the J/L/P
outputs computes each photon's replay weight and time from its record, then
returns a simulation that relaunches exactly those photons.
"""
from __future__ import annotations

import dataclasses
import struct
from typing import Optional

import numpy as np

from .sim import SimConfig, SimResult

R_C0 = np.float32(3.335640951981520e-12)    # src/mmc_const.h:45
HEADER = struct.Struct("<4s7If I f I i I 2i")


@dataclasses.dataclass
class MchFile:
    maxmedia: int
    detnum: int
    colcount: int
    totalphoton: int
    detected: int
    unitinmm: float
    normalizer: float
    records: np.ndarray                    # (savedphoton, colcount) float32
    seeds: Optional[np.ndarray] = None     # (savedphoton, 2) uint64


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def seeds_of(seed: int, ids) -> np.ndarray:
    """Initial xorshift128+ states of the given photon ids: the splitmix64
    rule of mmc_hip_photon_seeds, vectorized (uint64 arithmetic wraps)."""
    golden = np.uint64(0x9E3779B97F4A7C15)
    ids = np.asarray(ids, np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(int(seed) & (2**64 - 1)) + (np.uint64(2) * ids + np.uint64(1)) * golden
        s0, s1 = _mix64(z), _mix64(z + golden)
    s1 = np.where((s0 | s1) == 0, np.uint64(1), s1)
    return np.stack([s0, s1], 1)


def detected_seeds(cfg: SimConfig, res: SimResult) -> np.ndarray:
    """Initial xorshift128+ states of the photons behind res.detdata."""
    return seeds_of(cfg.seed, np.asarray(res.detphoton, np.int64))


def save_mch(path: str, prop: int, cfg: SimConfig, res: SimResult, unitinmm: float = 1.0,
             normalizer: float = 1.0, seeds: Optional[np.ndarray] = None):
    rec = np.ascontiguousarray(res.detdata, np.float32)
    n, col = rec.shape if rec.size else (0, res.detdata.shape[1] if res.detdata is not None else 0)
    seedbyte = 16 if seeds is not None else 0
    ndet = len(cfg.detpos) if cfg.detpos is not None else 0
    head = HEADER.pack(b"MCXH", 1, prop, ndet, col, int(cfg.nphoton), int(res.detected), n, unitinmm,
                       seedbyte, normalizer, 1, 1, 0, 0, 0)
    with open(path, "wb") as f:
        f.write(head)
        f.write(rec.tobytes())
        if seeds is not None:
            f.write(np.ascontiguousarray(seeds, np.uint64).reshape(n, 2).tobytes())


def load_mch(path: str) -> MchFile:
    raw = open(path, "rb").read()
    h = HEADER.unpack_from(raw, 0)
    if h[0] != b"MCXH":
        raise ValueError(f"{path}: not a detected-photon history file")
    maxmedia, detnum, colcount, total, detected, saved = h[2], h[3], h[4], h[5], h[6], h[7]
    unitinmm, seedbyte, normalizer = h[8], h[9], h[10]
    off = HEADER.size
    rec = np.frombuffer(raw, np.float32, saved * colcount, off).reshape(saved, colcount).copy()
    off += 4 * saved * colcount
    seeds = None
    if seedbyte:
        if seedbyte != 16:
            raise ValueError(f"{path}: {seedbyte}-byte seeds (xorshift128+ states are 16 bytes)")
        seeds = np.frombuffer(raw, np.uint64, 2 * saved, off).reshape(saved, 2).copy()
    return MchFile(maxmedia, detnum, colcount, total, detected, unitinmm, normalizer, rec, seeds)


def replay_config(cfg: SimConfig, mch: MchFile, med: np.ndarray, outputtype: str = "x",
                  replaydet: int = 0, detpattern=None, detparam1=None, detparam2=None) -> SimConfig:
    """The replay run of the saved photons of detector replaydet (0: all),
    with the replay weight w0 * prod_j exp(-mua_j * col_j * unitinmm) and
    time sum_j n_j * col_j * R_C0 over record columns j = 2..maxmedia+1 and
    media 1..maxmedia, as mesh_loadseedfile computes them in float.  With a
    detector pattern (detparam1.w x detparam2.w pixels over the plane at the
    first detector; records saved with exit positions) the starting weight is
    the pattern's value at each photon's exit pixel instead of w0
    (mesh_getdetweight, src/mmc_mesh.c:871-875, 2121-2138)."""
    if mch.seeds is None:
        raise ValueError("the history file holds no seeds (save them with the detected photons)")
    if mch.maxmedia != len(med) - 1:
        raise ValueError("the history file was generated with a different media setting")
    rec = mch.records
    keep = np.ones(len(rec), bool) if replaydet <= 0 else (rec[:, 0].astype(np.int64) == replaydet)
    rec = rec[keep]
    seeds = mch.seeds[keep]
    med = np.asarray(med, np.float32)
    rw = rec[:, -1].astype(np.float32).copy()
    if detpattern is not None:
        import ctypes as C
        from . import _lib
        r32 = np.ascontiguousarray(rec, np.float32)
        pat = np.ascontiguousarray(detpattern, np.float32).ravel()
        p1, p2 = np.asarray(detparam1, np.float32), np.asarray(detparam2, np.float32)
        org = np.asarray(cfg.detpos, np.float32).reshape(-1, 4)[0, :2].copy()
        if pat.size < int(p1[3]) * int(p2[3]):
            raise ValueError("detector pattern smaller than detparam1.w x detparam2.w")
        rc = _lib.lib().mmc_host_detweight(_lib.ptr(rw, C.c_float), _lib.ptr(r32, C.c_float), len(r32),
                                           r32.shape[1], _lib.ptr(org, C.c_float), _lib.ptr(p1, C.c_float),
                                           _lib.ptr(p2, C.c_float), _lib.ptr(pat, C.c_float))
        if rc != 0:
            raise ValueError("a photon's exit lies outside the detection plane (or records lack exit positions)")
    rt = np.zeros(len(rec), np.float32)
    unit = np.float32(mch.unitinmm)
    for j in range(2, mch.maxmedia + 2):
        rw *= np.exp(-med[j - 1, 0] * rec[:, j] * unit).astype(np.float32)
    for j in range(2, mch.maxmedia + 2):
        rt += (med[j - 1, 3] * rec[:, j] * R_C0).astype(np.float32)
    out = dataclasses.replace(cfg, nphoton=len(rec), photon_offset=0, outputtype=outputtype,
                              photonseed=seeds, replayweight=rw, replaytime=rt)
    if outputtype in ("j", "l", "p") or replaydet > 0:
        out.minenergy = 0.0
    return out
