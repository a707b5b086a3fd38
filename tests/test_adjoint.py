"""Adjoint Jacobians on the dual grid (mmc_hip_adjoint_grid; the reference's
mmc_adjoint_kernel / mmc_adjoint_dcoeff_kernel, src/mmc_core.cl:2228-2400,
scaled as src/mmc_cl_host.c:1436-1461).

The numpy restatement (oracle/adjoint.py) is pinned by closed forms: for
fields linear in x, y, z the second-order differences are exact, so
J_D = -u (a_s . a_d) at every voxel including the faces; for quadratic
fields the central and one-sided formulas are exact as well; J_mua of fields
constant over the grid is -u^3 cw_s cw_d; and an RF product obeys complex
multiplication.  The GPU entry point must equal the restatement.
"""
import numpy as np
import pytest

import adjoint as A

import mmc_amd as M
from mmc_amd import adjoint as MA
from mmc_amd import meshgen

NX, NY, NZ, G = 7, 5, 4, 3


def linear_fields(slopes, offsets):
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    f = [(a[0] * x + a[1] * y + a[2] * z + c) / G for a, c in zip(slopes, offsets)]
    return np.repeat(np.stack(f)[:, None], G, axis=1).astype(np.float32)   # [slot][gate][z][y][x]


def test_restatement_closed_forms():
    slopes = [(1.0, 2.0, -1.0), (0.5, -1.0, 2.0), (-2.0, 0.25, 1.0)]
    f = linear_fields(slopes, [30.0, 40.0, 50.0])          # 1 source, 2 detectors
    jm, jmi, jd, jdi = A.jacobians(f, None, 1, 2, unitinmm=2.0)
    assert jmi is None and jd.shape == (1, 2, NZ, NY, NX)
    for d in range(2):
        want = -2.0 * np.dot(slopes[0], slopes[1 + d])
        assert np.allclose(jd[0, d], want, rtol=1e-5, atol=1e-4)
    # constant fields: J_mua = -u^3 cw_s cw_d
    c = np.zeros_like(f)
    c[0], c[1], c[2] = 1.5, 2.0, -0.5
    jm, _, _, _ = A.jacobians(c, None, 1, 2, unitinmm=2.0)
    assert np.allclose(jm[0, 0], -8.0 * (3 * 1.5) * (3 * 2.0)) and np.allclose(jm[0, 1], -8.0 * 4.5 * -1.5)
    # quadratic along x: all three difference formulas are exact
    x = np.arange(NX, dtype=np.float32)
    q = np.zeros_like(f)
    q[0] = (x ** 2)[None, None, None, :] / G
    q[1] = x[None, None, None, :] / G
    _, _, jq, _ = A.jacobians(q, None, 1, 2, unitinmm=1.0)
    assert np.allclose(jq[0, 0], -(2 * x)[None, None, :] * 1.0, rtol=1e-6)
    # RF: complex products
    rng = np.random.default_rng(1)
    fr, fi = rng.random(f.shape, np.float32), rng.random(f.shape, np.float32)
    jm, jmi, jd, jdi = A.jacobians(fr, fi, 1, 2)
    cs = A.cw(fr) + 1j * A.cw(fi)
    assert np.allclose(jm[0, 1] + 1j * jmi[0, 1], -(cs[0] * cs[2]), rtol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("rf", [False, True])
def test_gpu_adjoint_equals_the_restatement(rf):
    rng = np.random.default_rng(7)
    ns, nd = 2, 3
    f = rng.random((ns + nd, G, NZ, NY, NX), np.float32)
    fi = rng.random(f.shape, np.float32) if rf else None
    j = MA.jacobian_grid(f, fi, ns, nd, unitinmm=0.5)
    jm, jmi, jd, jdi = A.jacobians(f, fi, ns, nd, unitinmm=0.5)
    assert np.allclose(j.mua, jm, rtol=1e-6, atol=0) and np.allclose(j.d, jd, rtol=1e-6, atol=1e-7)
    if rf:
        assert np.allclose(j.mua_im, jmi, rtol=1e-6, atol=0) and np.allclose(j.d_im, jdi, rtol=1e-6, atol=1e-7)


@pytest.mark.gpu
def test_gpu_adjoint_run_on_a_slab():
    # one source and two detectors on a layered slab: the detectors run as
    # sources into the medium; J_mua is negative where both fields reach
    arrays = meshgen.slab_layers(2.0, nx=40, ny=40, nz=20)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    src = M.SimConfig(nphoton=200000, seed=4, method="G", gridstep=2.0, srcpos=(14.0, 20.0, 0.0),
                      srcdir=(0.0, 0.0, 1.0), tstart=0.0, tend=5e-9, tstep=1e-9, outputtype="f")
    dets = MA.detector_sources(src, [(20.0, 20.0, 0.0), (26.0, 20.0, 0.0)], [(0, 0, 1), (0, 0, 1)])
    f, fi = MA.forward_fields(tm, [src], dets)
    assert fi is None and f.shape[0] == 3
    j = MA.jacobian_grid(f, None, 1, 2, unitinmm=1.0)
    jm, _, jd, _ = A.jacobians(f, None, 1, 2, unitinmm=1.0)
    assert np.allclose(j.mua, jm, rtol=1e-6, atol=0)
    assert np.all(j.mua <= 0) and j.mua.min() < 0
    # the sensitivity to the nearer detector is larger near the source
    assert j.mua[0, 0].sum() < j.mua[0, 1].sum()
