"""Nodal output under the branch-less Badouel tracer (-M S -C 1, the
reference's default basisorder 1).

The reference's CPU path deposits a third of each step's weight on each node
of the face the ray leaves through (src/mmc_raytrace.c:1674-1692); its GPU
backend instead runs element output and splits each element value in
quarters to its nodes (src/mmc_cl_host.c:973-1024).  The kernel follows the
CPU rule; mmc_host_elem_to_nodes gives the GPU one.  Both conserve the
deposited energy of each gate, which pins them against the element output of
the same photons."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import _lib
from mmc_amd._lib import ptr


def elem_to_nodes(ne, nn, elem, maxgate, srcnum, felem):
    out = np.zeros(maxgate * nn * srcnum, np.float64)
    rc = _lib.lib().mmc_host_elem_to_nodes(ne, nn, ptr(np.ascontiguousarray(elem, np.int32), C.c_int32), maxgate,
                                           srcnum, ptr(np.ascontiguousarray(felem, np.float64), C.c_double),
                                           ptr(out, C.c_double))
    assert rc == 0
    return out


def test_quarter_split_by_hand():
    # two tetrahedra sharing the face (2,3,4); two gates
    elem = np.array([[1, 2, 3, 4], [2, 3, 4, 5]], np.int32)
    felem = np.array([4.0, 8.0, 12.0, 0.0])          # [gate][elem]
    out = elem_to_nodes(2, 5, elem, 2, 1, felem).reshape(2, 5)
    assert np.allclose(out[0], [1, 3, 3, 3, 2])
    assert np.allclose(out[1], [3, 3, 3, 3, 0])
    # patterns interleave: [gate][elem][pattern] -> [gate][node][pattern]
    out2 = elem_to_nodes(2, 5, elem, 1, 2, np.array([4.0, 40.0, 8.0, 80.0])).reshape(5, 2)
    assert np.allclose(out2[:, 0], [1, 3, 3, 3, 2]) and np.allclose(out2[:, 1], 10 * out2[:, 0])


def slab_case(n):
    arrays = M.meshgen.slab_layers(4.0)
    cfg = M.SimConfig(nphoton=n, seed=31, method="S", srcpos=(50.3, 50.7, 0.0), srcdir=(0.0, 0.0, 1.0),
                      tend=2e-9, tstep=5e-10, outputtype="e")
    return arrays, cfg


def test_oracle_exit_face_thirds_conserve_each_gates_energy():
    arrays, cfg = slab_case(20000)
    om_e = O.OracleMesh(*arrays, nodal=False)
    om_n = O.OracleMesh(*arrays, nodal=True)
    seeds = M.photon_seeds(cfg.seed, 0, cfg.nphoton)
    fe = O.run(om_e, cfg, photonseed=seeds, records=(0, cfg.nphoton), math="portable")
    cn = M.SimConfig(**{**cfg.__dict__, "basisorder": 1})
    fn = O.run(om_n, cn, photonseed=seeds, records=(0, cfg.nphoton), math="portable")
    # the same photons: the output basis does not change the trajectories
    assert np.array_equal(fe.photons["p"], fn.photons["p"])
    se, sn = fe.field.sum(axis=1), fn.field.sum(axis=1)
    assert se.shape == sn.shape == (4,) and np.all(se > 0)
    assert np.allclose(sn, se, rtol=2e-6)
    # the quarter split conserves them too
    node, elem = arrays[0], arrays[1]
    q = elem_to_nodes(len(elem), len(node), elem, 4, 1, fe.field.ravel()).reshape(4, -1)
    assert np.allclose(q.sum(axis=1), se, rtol=1e-12)


@pytest.mark.gpu
def test_gpu_badouel_nodal_output_matches_the_oracle():
    arrays, cfg = slab_case(20000)
    cfg.basisorder = 1
    tm = M.TetMesh(*arrays).prepare(nodal=True)
    om = O.OracleMesh(*arrays, nodal=True)
    seeds = M.photon_seeds(cfg.seed, 0, cfg.nphoton)
    g = M.run(tm, cfg, records=(0, cfg.nphoton), tune=M.sim.tune_struct(compensated=1))
    o = O.run(om, cfg, photonseed=seeds, records=(0, cfg.nphoton), math="portable")
    assert np.array_equal(g.photons["eid"], o.photons["eid"])
    assert np.array_equal(g.photons["p"], o.photons["p"])
    assert g.field.shape == o.field.shape == (4, len(arrays[0]))
    assert np.linalg.norm(g.field - o.field) <= 1e-6 * np.linalg.norm(o.field)
    # energy per gate equals the element output of the same photons
    ce = M.SimConfig(**{**cfg.__dict__, "basisorder": 0})
    te = M.TetMesh(*arrays).prepare(nodal=False)
    ge = M.run(te, ce, tune=M.sim.tune_struct(compensated=1))
    assert np.allclose(g.field.sum(axis=1), ge.field.sum(axis=1), rtol=2e-6)
