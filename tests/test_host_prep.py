"""The product's host-side mesh preparation (include/mmc_hip_host.h) against the
oracle's restatement and the reference's fixtures, plus the C ABI surface."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as O
from conftest import ROOT, read_dat

import mmc_amd as M
from mmc_amd import _lib


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


MESHES = {
    "cube60_h4": lambda: M.meshgen.cube60(4.0),
    "slab_h4": lambda: M.meshgen.slab_layers(4.0),
    "jitter": lambda: M.meshgen.jittered_box(10, 8, 6, 1.0, amp=0.2, seed=5),
}


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("nodal", [False, True])
def test_prep_bitwise_equal_to_oracle(name, nodal):
    node, elem, et, med = MESHES[name]()
    tm = M.TetMesh(node, elem, et, med).prepare(nodal=nodal)
    om = O.OracleMesh(node, elem, et, med, nodal=nodal)
    for k in ("elem", "evol", "nvol", "facenb_raw", "facenb", "planes"):
        assert np.array_equal(bits(getattr(tm, k)), bits(getattr(om, k))), k
    assert tm.nf == om.nf
    pos = node.mean(axis=0)
    e, b = tm.locate(pos)
    e2, b2 = om.initelem(pos)
    assert e == e2 and np.array_equal(bits(b), bits(b2))


def test_prep_matches_reference_onecube(golden):
    d = os.path.join(golden, "onecube")
    node = read_dat(os.path.join(d, "node_onecube.dat"), 3).astype(np.float32)
    el = read_dat(os.path.join(d, "elem_onecube.dat"), 5).astype(np.int32)
    med = np.array([[0, 0, 1, 1], [0.005, 1.0101010101, 0.01, 1.0]], np.float32)
    tm = M.TetMesh(node, el[:, :4], el[:, 4], med).prepare()
    assert np.allclose(tm.evol, read_dat(os.path.join(d, "velem_onecube.dat"), 1)[:, 0], rtol=1e-6)
    assert np.allclose(tm.nvol, read_dat(os.path.join(d, "vnode_onecube.dat"), 1)[:, 0], rtol=1e-6)
    assert np.array_equal(tm.facenb_raw, read_dat(os.path.join(d, "facenb_onecube.dat"), 4, has_id=False))
    assert tm.nf == 12  # 6 cube faces x 2 triangles


def test_reff_and_normalize_match_oracle():
    L = _lib.lib()
    for n in (1.0, 1.33, 1.37, 1.45):
        assert L.mmc_host_reff(n, 1.0) == O.lib().orc_getreff(n, 1.0)
    node, elem, et, med = M.meshgen.cube60(6.0)
    for nodal in (False, True):
        tm = M.TetMesh(node, elem, et, med).prepare(nodal=nodal)
        om = O.OracleMesh(node, elem, et, med, nodal=nodal)
        cfg = M.SimConfig(nphoton=300, method="P" if nodal else "S", srcpos=(30.1, 30.2, 0), srcdir=(0, 0, 1),
                          tend=5e-9, tstep=1e-9)
        r = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, 300), math="portable")
        f1 = r.field.copy()
        k1 = O.lib().orc_normalize(C.byref(O.make_cfg(om, cfg)), C.byref(om.c), O.p(f1, C.c_double),
                                   C.c_float(r.absorbweight), C.c_float(r.launchweight))
        res = M.SimResult(field=r.field.copy(), fieldsum=0, launchweight=r.launchweight,
                          absorbweight=r.absorbweight, escapeweight=0, roulettegain=0, raytet=0, detected=0,
                          detdata=None, detphoton=None, photons=None, kernel_ms=0, total_ms=0)
        k2 = M.normalize_field(tm, cfg, res)
        assert k1 == k2
        assert np.array_equal(f1, res.field)


def test_normalized_element_field_conserves_absorbed_energy():
    # phi = raw/(V mua) * Eabs/(Etot sum raw) / dt  =>  sum phi V mua dt = Eabs/Etot
    node, elem, et, med = M.meshgen.slab_layers(4.0)
    om = O.OracleMesh(node, elem, et, med)
    cfg = M.SimConfig(nphoton=2000, method="S", srcpos=(50, 50, 0), srcdir=(0, 0, 1), tend=5e-9, tstep=1e-9)
    r = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), normalize=True)
    mua = om.med[om.etype, 0].astype(np.float64)
    total = (r.field * (om.evol * mua)[None, :]).sum() * cfg.gates()[3]
    assert abs(total - np.float32(r.absorbweight) / np.float32(r.launchweight)) < 1e-6


def declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(mmc_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header", ["mmc_hip.h", "mmc_hip_host.h"])
def test_library_exports_every_declared_function(header):
    path = os.path.join(ROOT, "include", header)
    names = declared_functions(path)
    assert len(names) >= 5
    lib = C.CDLL(_lib.LIBPATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert _lib.lib().mmc_hip_abi_version() == _lib.ABI_VERSION == 8


def test_photon_seeds_are_distinct_and_nonzero():
    s = M.photon_seeds(1648335518, 0, 100000)
    assert np.all((s[:, 0] | s[:, 1]) != 0)
    assert len(np.unique(s[:, 0])) == len(s)
    s2 = M.photon_seeds(1648335518, 50000, 10)
    assert np.array_equal(s2, s[50000:50010])
