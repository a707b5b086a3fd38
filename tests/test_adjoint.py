"""Adjoint Jacobians on the dual grid (mmc_hip_adjoint_grid; the reference's
mmc_adjoint_kernel / mmc_adjoint_dcoeff_kernel, src/mmc_core.cl:2228-2400,
scaled as src/mmc_cl_host.c:1436-1461).

The numpy restatement (oracle/adjoint.py) is pinned by closed forms: for
fields linear in x, y, z the second-order differences are exact, so
J_D = -u (a_s . a_d) at every voxel including the faces; for quadratic
fields the central and one-sided formulas are exact as well; J_mua of fields
constant over the grid is -u^3 cw_s cw_d; and an RF product obeys complex
multiplication.  The GPU entry point must equal the restatement.

Mesh mode (mmc_hip_adjoint_mesh; src/mmc_core.cl:2403-2649): the gradient
products <grad phi_i . grad phi_j> Ve are pinned on the unit right
tetrahedron and by zero row sums (the basis functions sum to one), the
host's cross-product form against the restatement's matrix inverse; the
element values by the mass matrix Ve/20 (1 + delta_ij) and by exact
stiffness integrals of linear fields, -Ve (a_s . a_d); the nodal output by
conservation (the nodes receive the elements' sums).
"""
import numpy as np
import pytest

import adjoint as A

import mmc_amd as M
from mmc_amd import adjoint as MA
from mmc_amd import meshgen

NX, NY, NZ, G = 7, 5, 4, 3


def small_mesh(n=4):
    node, elem = meshgen.t6_lattice(n, n, n, 1.5)
    node = node + np.random.default_rng(2).uniform(-0.1, 0.1, node.shape).astype(np.float32)
    etype = np.ones(len(elem), np.int32)
    med = np.array([[0, 0, 1, 1], [0.01, 1.0, 0.0, 1.37]], np.float32)
    return M.TetMesh(node, elem, etype, med).prepare(nodal=True)


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
    # (detectors of 1.5 mm radius launch as disks: the bottom layer is the
    # launch search's candidate list)
    arrays = meshgen.slab_layers(2.0, nx=40, ny=40, nz=20)
    srcelem = (np.where(meshgen.centroids(arrays[0], arrays[1])[:, 2] < 2.0)[0] + 1).astype(np.int32)
    tm = M.TetMesh(*arrays, srcelem=srcelem).prepare(nodal=False)
    src = M.SimConfig(nphoton=200000, seed=4, method="G", gridstep=2.0, srcpos=(14.0, 20.0, 0.0),
                      srcdir=(0.0, 0.0, 1.0), tstart=0.0, tend=5e-9, tstep=1e-9, outputtype="f")
    dets = MA.detector_sources(src, [(20.0, 20.0, 0.0, 1.5), (26.0, 20.0, 0.0, 1.5)], [(0, 0, 1), (0, 0, 1)])
    assert dets[0].srctype == "disk" and dets[0].srcparam1[0] == 1.5
    f, fi = MA.forward_fields(tm, [src], dets)
    assert fi is None and f.shape[0] == 3
    j = MA.jacobian_grid(f, None, 1, 2, unitinmm=1.0)
    jm, _, jd, _ = A.jacobians(f, None, 1, 2, unitinmm=1.0)
    assert np.allclose(j.mua, jm, rtol=1e-6, atol=0)
    assert np.all(j.mua <= 0) and j.mua.min() < 0
    # the sensitivity to the nearer detector is larger near the source
    assert j.mua[0, 0].sum() < j.mua[0, 1].sum()


def test_deldotdel_closed_forms():
    ref = np.array([0.5, -1, -1, -1, 1, 0, 0, 1, 0, 1], np.float64) / np.array([1, 6, 6, 6, 6, 1, 1, 6, 1, 6])
    node = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    assert np.allclose(A.deldotdel(node, [[1, 2, 3, 4]], [1 / 6]), ref, rtol=1e-12, atol=1e-15)
    tm = small_mesh()
    host = MA.deldotdel(tm)
    assert np.allclose(host, A.deldotdel(tm.node, tm.elem, tm.evol), rtol=1e-9, atol=1e-12)
    full = np.zeros((tm.ne, 4, 4))
    k = 0
    for i in range(4):
        for j in range(i, 4):
            full[:, i, j] = full[:, j, i] = host[:, k]
            k += 1
    assert np.abs(full.sum(2)).max() <= 1e-9 * np.abs(full).max()


def test_mesh_restatement_closed_forms():
    tm = small_mesh()
    ddd = A.deldotdel(tm.node, tm.elem, tm.evol).astype(np.float32)
    x = tm.node.astype(np.float64)
    a, b = np.array([0.3, -0.2, 0.5]), np.array([-0.1, 0.4, 0.2])
    f = np.stack([(x @ a + 0.5) / G, (x @ b + 0.25) / G]).astype(np.float32)   # 1 source, 1 detector
    f = np.repeat(f[:, None], G, axis=1)
    jm, _, jd, _ = A.mesh_element_values(f, None, tm.elem, tm.evol, ddd, 1, 1)
    # float32 sums of terms of both signs: cancellation costs a few digits
    assert np.allclose(jd[0, 0], -tm.evol * np.dot(a, b), rtol=1e-3, atol=1e-7)
    # mass matrix Ve/20 (1 + delta_ij)
    rng = np.random.default_rng(3)
    r = rng.random((2, G, tm.nn)).astype(np.float32)
    jm, _, _, _ = A.mesh_element_values(r, None, tm.elem, tm.evol, ddd, 1, 1)
    s, d = A.cw(r)[0][tm.elem - 1].astype(np.float64), A.cw(r)[1][tm.elem - 1].astype(np.float64)
    want = -tm.evol / 20 * ((s * d).sum(1) + s.sum(1) * d.sum(1))
    assert np.allclose(jm[0, 0], want, rtol=1e-5)
    # nodes receive the elements' sums; nodal approximation
    nod = A.to_nodes(jm, tm.elem, tm.nn)
    assert nod.sum() == pytest.approx(jm.sum(), rel=1e-5)
    e0 = tm.elem[0] - 1
    assert nod[0, 0, e0[0]] == pytest.approx(0.25 * jm[0, 0][(tm.elem - 1 == e0[0]).any(1)].sum(), rel=1e-5)
    ja, _ = A.mesh_nodal_approx(r, None, tm.nvol, 1, 1)
    assert np.allclose(ja[0, 0], -tm.nvol * A.cw(r)[0] * A.cw(r)[1], rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("rf", [False, True])
@pytest.mark.parametrize("mode", ["fem", "nodal", "elem"])
def test_gpu_mesh_adjoint_equals_the_restatement(mode, rf):
    tm = small_mesh()
    ns, nd = 2, 2
    rng = np.random.default_rng(11)
    f = rng.random((ns + nd, G, tm.nn), np.float32)
    fi = rng.random(f.shape, np.float32) if rf else None
    j = MA.jacobian_mesh(tm, f, fi, ns, nd, mode=mode)
    ddd = MA.deldotdel(tm).astype(np.float32)
    em, emi, ed, edi = A.mesh_element_values(f, fi, tm.elem, tm.evol, ddd, ns, nd)
    if mode == "elem":
        want = (em, emi, ed, edi)
    else:
        nodes = lambda v: None if v is None else A.to_nodes(v, tm.elem, tm.nn)
        want = [nodes(em), nodes(emi), nodes(ed), nodes(edi)]
        if mode == "nodal":
            want[0], want[1] = A.mesh_nodal_approx(f, fi, tm.nvol, ns, nd)
    for got, w in zip((j.mua, j.mua_im, j.d, j.d_im), want):
        assert (got is None) == (w is None)
        if w is not None:
            assert np.allclose(got, w, rtol=1e-6, atol=1e-6 * np.abs(w).max())


@pytest.mark.gpu
def test_gpu_mesh_adjoint_run_on_a_cube():
    node, elem = meshgen.t6_lattice(20, 20, 20, 3.0)
    etype = np.ones(len(elem), np.int32)
    med = np.array([[0, 0, 1, 1], [0.005, 1.0, 0.01, 1.37]], np.float32)
    tm = M.TetMesh(node, elem, etype, med).prepare(nodal=True)
    src = M.SimConfig(nphoton=200000, seed=5, method="P", srcpos=(24.5, 30.5, 0.0), srcdir=(0.0, 0.0, 1.0),
                      tstart=0.0, tend=5e-9, tstep=1e-9, outputtype="f")
    dets = MA.detector_sources(src, [(30.5, 30.5, 0.0), (36.5, 30.5, 0.0)], [(0, 0, 1), (0, 0, 1)])
    f, fi = MA.forward_fields(tm, [src], dets)
    assert fi is None and f.shape == (3, 5, tm.nn)
    j = MA.jacobian_mesh(tm, f, None, 1, 2)
    assert np.all(j.mua <= 0) and j.mua.min() < 0
    assert j.mua[0, 0].sum() < j.mua[0, 1].sum()


def test_adjoint_entry_points_refuse_bad_arguments():
    # checked before any device call, so this runs without a GPU
    import ctypes as C
    from mmc_amd import _lib
    L = _lib.lib()
    f = np.zeros(8, np.float32)
    dim = np.array([2, 2, 2], np.int32)
    out = np.zeros(8, np.float32)
    rc = L.mmc_hip_adjoint_grid(_lib.ptr(f, C.c_float), None, _lib.ptr(dim, C.c_int32), 1, 0, 1, C.c_float(1.0),
                                _lib.ptr(out, C.c_float), None)
    assert rc == -1 and "ns, nd" in _lib.last_error()
    elem = np.array([[1, 2, 3, 9]], np.int32)
    rc = L.mmc_hip_adjoint_mesh(_lib.ptr(f, C.c_float), None, 4, 1, _lib.ptr(elem, C.c_int32),
                                _lib.ptr(f, C.c_float), None, None, 1, 1, 1, 0, _lib.ptr(out, C.c_float), None)
    assert rc == -1 and "outside 1..nn" in _lib.last_error()
