"""Agreement tables between the deterministic FP32 functions and glibc.

The GPU kernel draws scattering lengths, azimuths and isotropic zeniths as the
FP32 functions of include/mmc_fp32math.h corrected by a per-generator-output
code from the library's tables (mmc_host_libm_table).  The oracle restates the
kernel's rule for applying a code and checks, over all 2^23 generator outputs,
that the corrected values equal glibc's logf / sincosf / acosf -- the
functions the reference CPU tracer calls (src/mmc_rand_common.h:50-51,
src/mmc_mesh.c:1662-1690, src/mmc_raytrace.c:2434-2480).
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O
from mmc_amd import _lib

NAMES = {0: "log", 1: "azimuth", 2: "isotropic", 3: "arcsine"}


class Esc(C.Structure):
    _fields_ = [("key", C.c_int32), ("v", C.c_uint32 * 3)]


def build_table(which):
    L = _lib.lib()
    n = L.mmc_host_libm_table_bytes(which)
    tab = np.zeros(n, np.uint8)
    esc = (Esc * 64)()
    ne = L.mmc_host_libm_table(which, tab.ctypes.data_as(C.c_void_p), C.cast(esc, C.c_void_p), 64)
    assert 0 <= ne <= 64
    flat = np.array([[e.key, *[np.int32(np.uint32(x)) for x in e.v]] for e in esc[:ne]], np.int32).reshape(-1, 4)
    return tab, np.ascontiguousarray(flat)


def verify(which, tab, esc):
    L = O.lib("portable")
    L.orc_libm_verify.restype = C.c_int64
    L.orc_libm_verify.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    return L.orc_libm_verify(which, tab.ctypes.data_as(C.c_void_p),
                             esc.ctypes.data_as(C.c_void_p) if len(esc) else None, len(esc))


@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_corrected_values_equal_glibc_on_every_generator_output(which):
    tab, esc = build_table(which)
    # the FP32 functions alone differ from glibc on a large share of inputs
    # (the codes are not all zero) ...
    assert np.count_nonzero(tab) > 10000, NAMES[which]
    # ... and with the codes applied they agree everywhere
    assert verify(which, tab, esc) == 0, NAMES[which]


def test_a_wrong_code_is_detected():
    tab, esc = build_table(0)
    k = int(np.flatnonzero(tab)[0])            # a byte holding a nonzero code
    bad = tab.copy()
    bad[k] ^= 0x3                              # flips the code of generator output 4k
    assert verify(0, bad, esc) >= 1
    tab1, esc1 = build_table(1)
    bad1 = tab1.copy()
    bad1[12345] = (bad1[12345] & 0xF3) | ((((bad1[12345] >> 2) & 3) ^ 1) << 2)   # cos code of output 24690
    assert verify(1, bad1, esc1) >= 1
