"""ctypes binding of libmmc_hip.so (the C ABI declared in include/mmc_hip.h and
include/mmc_hip_host.h).

The library is built in-tree by ``__graft_entry__.build()`` into
``mmc_amd/lib/libmmc_hip.so``.  Loading it fails loudly when it is missing:
there is no CPU fallback for the transport.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIBPATH = os.path.join(LIBDIR, "libmmc_hip.so")

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u64p = C.POINTER(C.c_uint64)


class MeshC(C.Structure):
    _fields_ = [
        ("nn", C.c_int32), ("ne", C.c_int32), ("prop", C.c_int32), ("nf", C.c_int32),
        ("node", f32p), ("elem", i32p), ("type", i32p), ("facenb", i32p),
        ("planes", f32p), ("med", f32p), ("srcelem", i32p), ("srcelemlen", C.c_int32),
        ("isextdet", C.c_int32),
    ]


class SimC(C.Structure):
    _fields_ = [
        ("nphoton", C.c_int64), ("seed", C.c_uint64), ("photon_offset", C.c_int64),
        ("method", C.c_int32), ("basisorder", C.c_int32), ("outputtype", C.c_int32),
        ("isreflect", C.c_int32), ("isspecular", C.c_int32), ("voidtime", C.c_int32),
        ("srctype", C.c_int32), ("srcnum", C.c_int32),
        ("e0", C.c_int32), ("bary0", C.c_float * 4),
        ("srcpos", C.c_float * 4), ("srcdir", C.c_float * 4),
        ("srcparam1", C.c_float * 4), ("srcparam2", C.c_float * 4),
        ("srcpattern", f32p),
        ("tstart", C.c_float), ("tend", C.c_float), ("tstep", C.c_float), ("maxgate", C.c_int32),
        ("nout", C.c_float), ("minenergy", C.c_float), ("roulettesize", C.c_float),
        ("issavedet", C.c_int32), ("issaveexit", C.c_int32),
        ("detnum", C.c_int32), ("detpos", f32p), ("maxdetphoton", C.c_int64),
        ("med", f32p),
        ("photonseed", C.POINTER(C.c_uint64)), ("replayweight", f32p), ("replaytime", f32p),
        ("gridstep", C.c_float), ("unitinmm", C.c_float), ("debuglevel", C.c_int32),
        ("omega", C.c_float), ("ismomentum", C.c_int32),
    ]


class PhotonC(C.Structure):
    _fields_ = [
        ("p", C.c_float * 3), ("w", C.c_float), ("v", C.c_float * 3), ("t", C.c_float),
        ("eid", C.c_int32), ("reason", C.c_int32), ("nraytet", C.c_int32), ("detid", C.c_int32),
        ("eabs", C.c_float), ("rgain", C.c_float), ("w0", C.c_float), ("nscat", C.c_float),
    ]


PHOTON_DTYPE = np.dtype([
    ("p", "<f4", 3), ("w", "<f4"), ("v", "<f4", 3), ("t", "<f4"),
    ("eid", "<i4"), ("reason", "<i4"), ("nraytet", "<i4"), ("detid", "<i4"),
    ("eabs", "<f4"), ("rgain", "<f4"), ("w0", "<f4"), ("nscat", "<f4"),
])
assert PHOTON_DTYPE.itemsize == C.sizeof(PhotonC) == 64


class ResultC(C.Structure):
    _fields_ = [
        ("field", f64p), ("fieldsum", C.c_double), ("launchweight", C.c_double), ("absorbweight", C.c_double),
        ("escapeweight", C.c_double), ("roulettegain", C.c_double), ("raytet", C.c_double),
        ("detdata", f32p), ("detphoton", i64p), ("detected", C.c_int64), ("reclen", C.c_int32),
        ("photons", C.c_void_p), ("recfirst", C.c_int64), ("reccount", C.c_int64),
        ("kernel_ms", C.c_double), ("total_ms", C.c_double), ("srcweight", f64p),
        ("traj", f32p), ("maxtraj", C.c_int64), ("ntraj", C.c_int64), ("field_im", f64p),
    ]


class TuneC(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("chunk", C.c_int32),
                ("compensated", C.c_int32), ("replicas", C.c_int32)]


_lib = None

# exported symbols and their signatures: (restype, argtypes)
SIGNATURES = {
    "mmc_hip_abi_version": (C.c_int32, []),
    "mmc_hip_last_error": (C.c_char_p, []),
    "mmc_hip_device_info": (C.c_int32, [C.c_char_p, C.c_int32, i64p]),
    "mmc_hip_photon_seeds": (None, [C.c_uint64, C.c_int64, C.c_int64, u64p]),
    "mmc_hip_run": (C.c_int32, [C.POINTER(MeshC), C.POINTER(SimC), C.POINTER(TuneC), C.POINTER(ResultC)]),
    "mmc_hip_mesh_upload": (C.c_int32, [C.POINTER(MeshC), C.POINTER(C.c_void_p)]),
    "mmc_hip_mesh_release": (C.c_int32, [C.c_void_p]),
    "mmc_hip_mesh_bytes": (C.c_int64, [C.c_void_p]),
    "mmc_hip_run_batch": (C.c_int32, [C.c_void_p, C.POINTER(SimC), C.c_int32, C.POINTER(TuneC),
                                      C.POINTER(ResultC)]),
    "mmc_hip_sim_output_bytes": (C.c_int64, [C.c_void_p, C.POINTER(SimC)]),
    "mmc_hip_sim_grid": (C.c_int32, [C.c_void_p, C.POINTER(SimC), f32p, i32p]),
    "mmc_hip_adjoint_grid": (C.c_int32, [f32p, f32p, i32p, C.c_int32, C.c_int32, C.c_int32, C.c_float, f32p,
                                         f32p]),
    "mmc_hip_adjoint_mesh": (C.c_int32, [f32p, f32p, C.c_int32, C.c_int32, i32p, f32p, f32p, f32p, C.c_int32,
                                         C.c_int32, C.c_int32, C.c_int32, f32p, f32p]),
    "mmc_hip_arena_create": (C.c_int32, [C.c_uint64, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "mmc_hip_arena_release": (C.c_int32, [C.c_void_p]),
    "mmc_hip_run_chunks": (C.c_int32, [C.c_void_p, C.POINTER(SimC), C.c_int32, C.c_void_p, u64p,
                                       C.POINTER(TuneC), C.POINTER(ResultC)]),
    "mmc_hip_arena_collect": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_uint64, f64p, f64p]),
    "mmc_hip_run_chunks_noalloc": (C.c_int32, [C.c_void_p, C.POINTER(SimC), C.c_int32, C.c_void_p, u64p,
                                               C.POINTER(TuneC), C.POINTER(ResultC)]),
    "mmc_hip_slot_create": (C.c_int32, [C.POINTER(C.c_void_p)]),
    "mmc_hip_slot_release": (C.c_int32, [C.c_void_p]),
    "mmc_hip_run_chunks_async": (C.c_int32, [C.c_void_p, C.POINTER(SimC), C.c_int32, C.c_void_p, u64p,
                                             C.POINTER(TuneC), C.POINTER(ResultC), C.c_void_p]),
    "mmc_hip_slot_query": (C.c_int32, [C.c_void_p]),
    "mmc_hip_slot_wait": (C.c_int32, [C.c_void_p, C.POINTER(ResultC)]),
    "mmc_host_volume": (None, [C.c_int32, f32p, C.c_int32, i32p, i32p, f32p, f32p]),
    "mmc_host_facenb": (C.c_int32, [C.c_int32, C.c_int32, i32p, i32p]),
    "mmc_host_planes": (None, [C.c_int32, f32p, i32p, f32p]),
    "mmc_host_barycentric": (C.c_int32, [f32p, i32p, C.c_int32, f32p, f32p]),
    "mmc_host_initelem": (C.c_int32, [C.c_int32, f32p, i32p, f32p, f32p]),
    "mmc_host_reff": (C.c_double, [C.c_double, C.c_double]),
    "mmc_host_surface_nvol": (None, [C.c_int32, C.c_int32, i32p, i32p, i32p, f32p, C.c_int32,
                                     C.c_int32, f32p]),
    "mmc_host_number_exterior": (C.c_int32, [C.c_int32, i32p]),
    "mmc_host_detelem": (C.c_int32, [C.c_int32, i32p, C.c_int32]),
    "mmc_host_elem_to_nodes": (C.c_int32, [C.c_int32, C.c_int32, i32p, C.c_int32, C.c_int32, f64p, f64p]),
    "mmc_host_grid": (C.c_int32, [f32p, C.c_int32, C.c_float, f32p, i32p]),
    "mmc_host_srcelem": (C.c_int32, [C.c_int32, i32p, i32p]),
    "mmc_host_detimage": (C.c_int32, [f32p, f32p, C.c_int64, C.c_int32, C.c_int32, f32p, C.c_float, C.c_float,
                                      C.c_int32, f32p, f32p, f32p]),
    "mmc_host_detweight": (C.c_int32, [f32p, f32p, C.c_int64, C.c_int32, f32p, f32p, f32p, f32p]),
    "mmc_host_deldotdel": (C.c_int32, [f32p, C.c_int32, i32p, C.c_int32, f32p, f64p]),
    "mmc_host_normalize_grid": (C.c_double, [C.POINTER(SimC), C.c_int64, f64p, C.c_double, C.c_double,
                                             C.c_int32]),
    "mmc_host_normalize": (C.c_double, [C.POINTER(SimC), C.c_int32, C.c_int32, i32p, i32p, f32p,
                                        f32p, f32p, f64p, C.c_double, C.c_double]),
    "mmc_host_normalize_pair": (C.c_double, [C.POINTER(SimC), C.c_int32, C.c_int32, i32p, i32p, f32p,
                                             f32p, f32p, f64p, C.c_double, C.c_double, C.c_int32]),
    "mmc_host_save_dat": (C.c_int32, [C.c_char_p, f64p, C.c_int32, C.c_int32]),
    "mmc_host_save_bin": (C.c_int32, [C.c_char_p, f64p, C.c_int64]),
    "mmc_host_save_mch": (C.c_int32, [C.c_char_p, C.c_void_p, f32p, C.c_int64]),
    "mmc_host_save_field": (C.c_int32, [C.c_char_p, C.c_int32, f64p, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_float, C.c_int32, C.c_int32]),
    "mmc_host_save_grid": (C.c_int32, [C.c_char_p, C.c_int32, f64p, i32p, C.c_float, C.c_int32, C.c_int32,
                                       C.c_float, C.c_int32, C.c_int32]),
    "mmc_host_format_id": (C.c_int32, [C.c_char_p]),
    "mmc_host_libm_table_bytes": (C.c_int64, [C.c_int32]),
    "mmc_host_libm_table": (C.c_int32, [C.c_int32, C.c_void_p, C.c_void_p]),
    "mmc_host_libm_nval": (C.c_int32, [C.c_int32]),
}


ABI_VERSION = 8        # MMC_HIP_ABI_VERSION of include/mmc_hip.h


def lib():
    """Load libmmc_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: run __graft_entry__.build() first "
                               "(the photon transport has no CPU fallback)")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.mmc_hip_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIBPATH} has ABI {L.mmc_hip_abi_version()}, this binding expects "
                               f"{ABI_VERSION}: rebuild with __graft_entry__.build()")
        _lib = L
    return _lib


def ptr(a, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


def last_error() -> str:
    return lib().mmc_hip_last_error().decode(errors="replace")


def check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
