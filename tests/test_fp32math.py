"""Accuracy of the deterministic FP32 libm (include/mmc_fp32math.h) against
double precision, and the reference's double-precision azimuth
(float)(2*pi*(double)u) (src/mmc_mesh.c:1666) for every random number the
generator can produce."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "_build")
SRC = r"""
#include "mmc_fp32math.h"
void v_exp(const float* x, float* y, long n) { for (long i = 0; i < n; i++) y[i] = mmc_expf(x[i]); }
void v_log(const float* x, float* y, long n) { for (long i = 0; i < n; i++) y[i] = mmc_logf(x[i]); }
void v_sincos(const float* x, float* s, float* c, long n) { for (long i = 0; i < n; i++) mmc_sincosf(x[i], s + i, c + i); }
void v_acos(const float* x, float* y, long n) { for (long i = 0; i < n; i++) y[i] = mmc_acosf(x[i]); }
void v_twopi(const float* u, float* y, long n) { for (long i = 0; i < n; i++) y[i] = mmc_twopi(u[i]); }
"""


@pytest.fixture(scope="module")
def mlib():
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(BUILD, "mathprobe.c")
    so = os.path.join(BUILD, "libmathprobe.so")
    with open(src, "w") as f:
        f.write(SRC)
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-mfma", "-I", os.path.join(ROOT, "include"),
                    src, "-o", so], check=True)
    return C.CDLL(so)


def call(L, name, *arrays):
    args = [a.ctypes.data_as(C.c_void_p) for a in arrays] + [C.c_long(len(arrays[0]))]
    getattr(L, name)(*args)


def ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    ulp = np.abs(np.nextafter(ref32, np.float32(np.inf)) - ref32).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / ulp


def test_exp(mlib):
    x = np.linspace(-103.0, 88.0, 2_000_001).astype(np.float32)
    y = np.zeros_like(x)
    call(mlib, "v_exp", x, y)
    ref = np.exp(x.astype(np.float64))
    normal = ref > 1.2e-38
    assert ulp_err(y[normal], ref[normal]).max() <= 1.5
    assert np.all(np.abs(y[~normal] - ref[~normal]) <= 2e-45)
    z = np.array([0.0, -0.0], np.float32)
    w = np.zeros_like(z)
    call(mlib, "v_exp", z, w)
    assert np.all(w == 1.0)


def test_log(mlib):
    x = np.concatenate([np.geomspace(1e-6, 1.0 + 1e-6, 2_000_000), np.geomspace(1e-30, 1e30, 500_000)]).astype(np.float32)
    y = np.zeros_like(x)
    call(mlib, "v_log", x, y)
    ref = np.log(x.astype(np.float64))
    nz = np.abs(ref) > 1e-30
    assert ulp_err(y[nz], ref[nz]).max() <= 1.5
    one = np.array([1.0], np.float32)
    r = np.zeros_like(one)
    call(mlib, "v_log", one, r)
    assert r[0] == 0.0


def test_sincos(mlib):
    x = np.linspace(0.0, 2 * np.pi, 2_000_001).astype(np.float32)
    s, c = np.zeros_like(x), np.zeros_like(x)
    call(mlib, "v_sincos", x, s, c)
    xd = x.astype(np.float64)
    assert np.abs(s - np.sin(xd)).max() < 1.5e-7
    assert np.abs(c - np.cos(xd)).max() < 1.5e-7
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 5e-7


def test_acos(mlib):
    x = np.linspace(-1.0, 1.0, 1_000_001).astype(np.float32)
    y = np.zeros_like(x)
    call(mlib, "v_acos", x, y)
    assert np.abs(y - np.arccos(x.astype(np.float64))).max() < 5e-7


def test_twopi_matches_double_product_exhaustively(mlib):
    m = np.arange(1 << 23, dtype=np.uint32)
    u = (m.astype(np.float64) * 2.0 ** -23).astype(np.float32)     # every generator output
    y = np.zeros(len(m), np.float32)
    call(mlib, "v_twopi", u, y)
    ref = ((2.0 * np.pi) * u.astype(np.float64)).astype(np.float32)
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))
    assert y[0] == 0.0 and y[-1] < np.float32(2 * np.pi)
