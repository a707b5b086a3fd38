"""Pins of the oracle's geometry and optics against the reference's own mesh
fixtures (examples/onecube, examples/validation) and against what the
mathematics fixes: brute-force ray/plane intersection in double precision,
the textbook Fresnel formula, the Henyey-Greenstein moments and CDF, and the
effective reflection coefficient."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoracle as O
from conftest import read_dat

import mmc_amd as M


def onecube(golden):
    d = os.path.join(golden, "onecube")
    node = read_dat(os.path.join(d, "node_onecube.dat"), 3).astype(np.float32)
    el = read_dat(os.path.join(d, "elem_onecube.dat"), 5).astype(np.int32)
    return node, el[:, :4].copy(), el[:, 4].copy(), d


def test_volumes_and_neighbours_match_reference_fixture(golden):
    node, elem, et, d = onecube(golden)
    med = np.array([[0, 0, 1, 1], [0.005, 1.0101, 0.01, 1.0]], np.float32)
    om = O.OracleMesh(node, elem, et, med)
    velem = read_dat(os.path.join(d, "velem_onecube.dat"), 1)[:, 0]
    vnode = read_dat(os.path.join(d, "vnode_onecube.dat"), 1)[:, 0]
    fnb = read_dat(os.path.join(d, "facenb_onecube.dat"), 4, has_id=False).astype(np.int32)
    assert np.allclose(om.evol, velem, rtol=1e-6)
    assert np.allclose(om.nvol, vnode, rtol=1e-6)
    assert np.array_equal(om.facenb_raw, fnb)
    # cube2 (validation example): 60 mm cube, same connectivity
    d2 = os.path.join(golden, "cube2")
    node2 = read_dat(os.path.join(d2, "node_cube2.dat"), 3).astype(np.float32)
    el2 = read_dat(os.path.join(d2, "elem_cube2.dat"), 5).astype(np.int32)
    om2 = O.OracleMesh(node2, el2[:, :4], el2[:, 4], med)
    assert np.allclose(om2.evol, read_dat(os.path.join(d2, "velem_cube2.dat"), 1)[:, 0], rtol=1e-6)
    assert np.array_equal(om2.facenb_raw, read_dat(os.path.join(d2, "facenb_cube2.dat"), 4, has_id=False).astype(np.int32))


def test_planes_and_plucker_coordinates():
    node, elem, et, med = M.meshgen.jittered_box(6, 5, 4, 1.0, amp=0.2, seed=3)
    om = O.OracleMesh(node, elem, et, med)
    P = om.planes.reshape(-1, 4, 4)        # [elem][nx,ny,nz,d][face]
    n = np.transpose(P[:, :3, :], (0, 2, 1)).astype(np.float64)   # [elem][face][xyz]
    d = P[:, 3, :].astype(np.float64)
    assert np.allclose(np.linalg.norm(n, axis=2), 1.0, atol=1e-6)
    out = np.array([[0, 3, 1], [3, 2, 1], [0, 2, 3], [0, 1, 2]])
    opp = np.array([2, 0, 1, 3])
    X = om.node[om.elem - 1].astype(np.float64)        # [elem][node][xyz]
    for f in range(4):
        for k in range(3):
            assert np.allclose(np.einsum("ij,ij->i", n[:, f], X[:, out[f, k]]), d[:, f], atol=1e-4)
        # outward: the opposite node lies behind the plane
        assert np.all(np.einsum("ij,ij->i", n[:, f], X[:, opp[f]]) < d[:, f])
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    pd, pm = om.pd.reshape(-1, 6, 3), om.pm.reshape(-1, 6, 3)
    for j, (a, b) in enumerate(pairs):
        assert np.allclose(pd[:, j], X[:, b] - X[:, a], atol=1e-5)
        assert np.allclose(pm[:, j], np.cross(X[:, a], X[:, b]), atol=1e-3)


def brute_exit(X, p, v):
    """Exit face (facenb column) and distance of the ray p + t v from the
    tetrahedron X[4,3], in double precision."""
    out = np.array([[0, 3, 1], [3, 2, 1], [0, 2, 3], [0, 1, 2]])
    faceorder = [1, 3, 2, 0]
    best, bf = np.inf, -1
    for f in range(4):
        A, B, Cc = X[out[f, 0]], X[out[f, 1]], X[out[f, 2]]
        n = np.cross(B - A, Cc - A)
        s = n @ v
        if s > 0:
            t = (n @ (A - p)) / s
            if t < best:
                best, bf = t, faceorder[f]
    return bf, best


@pytest.mark.parametrize("method", [0, 3])
def test_ray_tet_exit_matches_brute_force(method):
    node, elem, et, med = M.meshgen.jittered_box(5, 5, 5, 1.0, amp=0.25, seed=11)
    om = O.OracleMesh(node, elem, et, med, nodal=(method == 0))
    rng = np.random.default_rng(1)
    L = O.lib()
    out = np.zeros(8, np.float32)
    agree = 0
    ntest = 3000
    for _ in range(ntest):
        e = int(rng.integers(1, om.ne + 1))
        X = om.node[om.elem[e - 1] - 1].astype(np.float64)
        w = rng.dirichlet(np.ones(4))
        p = (w @ X).astype(np.float32)
        v = rng.normal(size=3)
        v = (v / np.linalg.norm(v)).astype(np.float32)
        L.orc_probe_raytet(C.byref(om.c), method, e, O.p(p, C.c_float), O.p(v, C.c_float), C.c_float(1e6),
                           O.p(out, C.c_float))
        bf, bt = brute_exit(X, p.astype(np.float64), v.astype(np.float64))
        if out[6] == 1 and int(out[0]) == bf:
            agree += 1
            assert abs(out[2] - bt) < 1e-4
            assert np.allclose(out[3:6], p + bt * v, atol=1e-4)
    assert agree >= ntest - 3        # only rays grazing an edge may differ


def fresnel(n1, n2, cosi):
    sint2 = (n1 / n2) ** 2 * (1 - cosi ** 2)
    if sint2 >= 1:
        return 1.0
    cost = np.sqrt(1 - sint2)
    rs = ((n1 * cosi - n2 * cost) / (n1 * cosi + n2 * cost)) ** 2
    rp = ((n1 * cost - n2 * cosi) / (n1 * cost + n2 * cosi)) ** 2
    return 0.5 * (rs + rp)


def test_fresnel_reflectance_and_total_internal_reflection():
    # two-element mesh: element 1 (n=1.37) and an exterior face; reflectray returns
    # the unpolarized reflectance when called with eid == oldeid
    node, elem, et, med = M.meshgen.cube60(30.0)
    med = med.copy()
    om = O.OracleMesh(node, elem, et, med)
    L = O.lib()
    st = (C.c_uint64 * 2)(1, 2)
    P = om.planes[0].reshape(4, 4)
    for face in range(4):
        f = [3, 0, 2, 1][face]
        n = P[:3, f].astype(np.float64)
        for ang in np.linspace(0.0, 1.5, 16):
            t = np.cross(n, [0.3, 0.5, 0.8])
            t /= np.linalg.norm(t)
            v = (np.cos(ang) * n + np.sin(ang) * t).astype(np.float32)
            eid = C.c_int(1)
            R = L.orc_reflect(C.byref(om.c), 3, C.c_float(1.0), 1, 0, O.p(v.copy(), C.c_float), 1,
                              C.byref(eid), face, st)
            # eid == oldeid is the specular case: from the outside (n=1) into n=1.37
            assert abs(R - fresnel(1.0, 1.37, np.cos(ang))) < 2e-6
    # exterior: n1 = 1.37 -> n2 = nout; test against the textbook formula
    for nout in (1.0, 1.33, 1.6):
        for ang in np.linspace(0.0, 1.4, 15):
            f = 0
            face = [1, 3, 2, 0][f]
            n = P[:3, f].astype(np.float64)
            t = np.cross(n, [0.3, 0.5, 0.8])
            t /= np.linalg.norm(t)
            v = (np.cos(ang) * n + np.sin(ang) * t).astype(np.float32)
            eid = C.c_int(-1)
            counts = 0
            trials = 4000
            for _ in range(trials):
                eid.value = -1
                vv = v.copy()
                L.orc_reflect(C.byref(om.c), 3, C.c_float(nout), 1, 0, O.p(vv, C.c_float), 1, C.byref(eid), face, st)
                counts += eid.value == 1
            Rref = fresnel(1.37, nout, np.cos(ang))
            sigma = np.sqrt(max(Rref * (1 - Rref), 1e-4) / trials)
            assert abs(counts / trials - Rref) < 5 * sigma + 1e-3
            if Rref == 1.0:     # total internal reflection mirrors the direction
                vv = v.copy()
                eid.value = -1
                L.orc_reflect(C.byref(om.c), 3, C.c_float(nout), 1, 0, O.p(vv, C.c_float), 1, C.byref(eid), face, st)
                assert np.allclose(vv, v - 2 * (v @ n) * n, atol=1e-5)


@pytest.mark.parametrize("math", ["glibc", "portable"])
@pytest.mark.parametrize("g", [0.0, 0.01, 0.5, 0.9])
def test_henyey_greenstein(g, math):
    L = O.lib(math)
    st = (C.c_uint64 * 2)(123, 456)
    N = 100000
    cos = np.zeros(N)
    d0 = np.array([0.3, -0.4, np.sqrt(1 - 0.25)], np.float32)
    d0 /= np.linalg.norm(d0)
    for i in range(N):
        d = d0.copy()
        L.orc_scatter(C.c_float(g), O.p(d, C.c_float), st)
        assert abs(np.linalg.norm(d) - 1) < 2e-6
        cos[i] = d.astype(np.float64) @ d0
    assert abs(cos.mean() - g) < 5 * np.sqrt((1 - g * g) / 3 / N) + 2e-3
    xs = np.sort(cos)
    if g > 1e-6:
        cdf = (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * xs) - 1 / (1 + g))
    else:
        cdf = (xs + 1) / 2
    assert np.max(np.abs(cdf - np.arange(1, N + 1) / N)) < 1.63 / np.sqrt(N) + 1e-3


def test_effective_reflection_coefficient():
    L = O.lib()
    assert abs(L.orc_getreff(1.0, 1.0)) < 1e-12
    # published values (Haskell et al. 1994, n_out = 1): n=1.33 -> 0.431, n=1.4 -> 0.493
    assert abs(L.orc_getreff(1.33, 1.0) - 0.431) < 0.01
    assert abs(L.orc_getreff(1.4, 1.0) - 0.493) < 0.01


def _exact_containing(node, elem, p):
    """Elements containing p by the signs of the four sub-volumes computed in
    double precision (the first-match rule of mesh_initelem over them):
    returns a (ne,) array of the smallest signed sub-volume per element."""
    v = node[elem - 1].astype(np.float64)            # (ne, 4, 3)
    q = np.asarray(p, np.float64)
    # the sub-volume opposite node i: replace node i by p in the determinant
    def vol(a, b, c, d):
        return np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    full = vol(v[:, 0], v[:, 1], v[:, 2], v[:, 3])
    sub = []
    for i in range(4):
        w = v.copy()
        w[:, i] = q
        sub.append(vol(w[:, 0], w[:, 1], w[:, 2], w[:, 3]) * np.sign(full))
    return np.min(np.stack(sub, 1), axis=1)


def test_initelem_first_match_and_barycentrics_by_brute_force():
    """orc_initelem (mesh_initelem, src/mmc_mesh.c:1060-1090) against an
    exhaustive double-precision search: (1) points inside a jittered mesh lie
    in exactly one element by a clear margin, which must be the one found;
    (2) lattice nodes and edge midpoints of an unjittered lattice (all
    products exact in FP32) lie in several elements, and the first match in
    element order must be the lowest-index element containing them;
    barycentrics are non-negative, sum to 1 and reconstruct the point."""
    rng = np.random.default_rng(5)
    # (1) interior points of a jittered mesh
    node, elem = M.meshgen.t6_lattice(6, 6, 4, 1.0)
    node = M.meshgen.jitter_interior(node, 0.25, 3, (0.0, 0.0, 0.0), (6.0, 6.0, 4.0))
    et = np.ones(len(elem), np.int32)
    med = np.array([[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37]], np.float32)
    om = O.OracleMesh(node, elem, et, med, nodal=False)
    checked = 0
    for _ in range(300):
        p = rng.uniform((0.1, 0.1, 0.1), (5.9, 5.9, 3.9)).astype(np.float32)
        m = _exact_containing(node, elem, p)
        inside = np.flatnonzero(m >= -1e-9)
        if len(inside) != 1 or m[inside[0]] < 1e-4:
            continue                                 # near a face: the FP32 test may differ
        e, b = om.initelem(p)
        assert e == inside[0] + 1
        assert np.all(b >= 0) and abs(b.sum() - 1) < 1e-5
        # barycentrics follow the element's nodes after the orientation fix of
        # mesh_getvolume (the oracle mesh's element order)
        rec = (b[:, None] * node[om.elem[e - 1] - 1]).sum(axis=0)
        assert np.abs(rec - p).max() < 1e-4
        checked += 1
    assert checked > 200
    # (2) shared lattice points: first match in element order
    node, elem = M.meshgen.t6_lattice(4, 4, 3, 1.0)
    et = np.ones(len(elem), np.int32)
    om = O.OracleMesh(node, elem, et, med, nodal=False)
    pts = [node[i] for i in rng.choice(len(node), 40, replace=False)]
    pts += [0.5 * (node[elem[k, 0] - 1] + node[elem[k, j] - 1]) for k in rng.choice(len(elem), 40) for j in (1, 3)]
    nshared = 0
    for p in pts:
        p = np.asarray(p, np.float32)
        m = _exact_containing(node, elem, p)
        first = int(np.flatnonzero(m >= 0)[0]) + 1
        e, b = om.initelem(p)
        assert e == first
        assert np.all(b >= 0) and abs(b.sum() - 1) < 1e-6
        nshared += (m >= 0).sum() > 1
    assert nshared > 40
