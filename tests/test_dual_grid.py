"""Dual-grid MMC (DMMC, "-M G"): branch-less Badouel tracing with the
deposits on a voxel grid over the mesh's bounding box.

The oracle restates the CPU deposit (src/mmc_raytrace.c:1615-1673) and the
grid setup (mesh_createdualmesh, src/mmc_mesh.c:349-381).  It is pinned by
closed forms the algorithm fixes: the grid dimensions of a known box; a
non-scattering beam along a lattice column, whose every step lies inside one
voxel layer, so voxel k holds exactly the Beer-Lambert loss over [k, k+1]
and nothing lands outside the beam's column; the energy total equals the
element-output total of the same photons (the grid only moves deposits);
and the normalization 1/(Etotal unitinmm^3) (src/mmc_mesh.c:2205-2206).
GPU tests compare the HIP kernel with that oracle on the same photons.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import _lib, meshgen
from mmc_amd.replay import seeds_of

MU = 0.1


def beam_box():
    """10 x 10 x 20 mm homogeneous box of 1 mm cells (6 tets each),
    mua = 0.1/mm, no scattering, index matched to the outside."""
    node, elem = meshgen.t6_lattice(10, 10, 20, 1.0)
    etype = np.ones(len(elem), np.int32)
    med = np.array([[0, 0, 1, 1], [MU, 0.0, 0.0, 1.0]], np.float32)
    return node, elem, etype, med


def beam_sim(n=64, out="e", step=1.0):
    # x = 5.25, y = 3.75: inside voxel column (5, 3), away from the lattice's
    # diagonal faces' intersections with the cell faces
    return M.SimConfig(nphoton=n, seed=11, method="G", gridstep=step, srcpos=(5.25, 3.75, 0.0),
                       srcdir=(0.0, 0.0, 1.0), tstart=0.0, tend=1e-8, tstep=1e-8, outputtype=out,
                       isreflect=0, nout=1.0, minenergy=0.0)


def test_grid_dimensions_of_a_known_box():
    node = np.array([[0, 0, 0], [10, 0, 0], [0, 6, 0], [0, 0, 3.5]], np.float32)
    nmin = np.zeros(3, np.float32)
    dim = np.zeros(3, np.int32)
    L = _lib.lib()
    assert L.mmc_host_grid(_lib.ptr(node, C.c_float), 4, C.c_float(1.0), _lib.ptr(nmin, C.c_float),
                           _lib.ptr(dim, C.c_int32)) == 0
    # extents 10, 6, 3.5 widened by 2e-6: int(extent / 1) + 1 voxels
    assert list(dim) == [11, 7, 4]
    assert np.all(nmin == np.float32(0) - np.float32(1e-6))
    assert L.mmc_host_grid(_lib.ptr(node, C.c_float), 4, C.c_float(0.3), _lib.ptr(nmin, C.c_float),
                           _lib.ptr(dim, C.c_int32)) == 0
    assert list(dim) == [34, 21, 12]
    assert L.mmc_host_grid(_lib.ptr(node, C.c_float), 4, C.c_float(0.0), _lib.ptr(nmin, C.c_float),
                           _lib.ptr(dim, C.c_int32)) == -1
    # the oracle's restatement agrees
    om = O.OracleMesh(*beam_box(), nodal=False, isreflect=0)
    omin, odim = O.grid(om, 1.0)
    assert odim == [11, 11, 21]
    cfg = beam_sim()
    tm_min, tm_dim = cfg.grid(M.TetMesh(*beam_box()))
    assert list(tm_dim) == odim and np.allclose(tm_min, omin, rtol=0, atol=0)


def test_straight_beam_deposits_the_beer_lambert_loss_per_voxel():
    om = O.OracleMesh(*beam_box(), nodal=False, isreflect=0)
    cfg = beam_sim(n=64)
    o = O.run(om, cfg, photonseed=seeds_of(cfg.seed, np.arange(cfg.nphoton)), math="portable")
    nmin, dim = O.grid(om, 1.0)
    f = o.field.reshape(1, dim[2], dim[1], dim[0])[0]       # [z][y][x]
    col = f[:, 3, 5]
    assert np.count_nonzero(f) == np.count_nonzero(col) == 20
    k = np.arange(20)
    want = cfg.nphoton * (np.exp(-MU * k) - np.exp(-MU * (k + 1)))
    assert np.allclose(col[:20], want, rtol=2e-5, atol=0)
    assert col[20] == 0
    assert o.absorbweight == pytest.approx(cfg.nphoton * (1 - np.exp(-MU * 20)), rel=1e-5)
    # fluence output deposits the loss divided by mua; normalization is
    # 1 / (Etotal unitinmm^3)
    cfg.outputtype, cfg.unitinmm = "f", 2.0
    o2 = O.run(om, cfg, photonseed=seeds_of(cfg.seed, np.arange(cfg.nphoton)), math="portable", normalize=True)
    assert o2.normalizer == pytest.approx(1.0 / (cfg.nphoton * 8.0), rel=1e-7)
    col2 = o2.field.reshape(dim[2], dim[1], dim[0])[:, 3, 5]
    assert np.allclose(col2[:20], want / MU / (cfg.nphoton * 8.0), rtol=5e-5)


def test_half_voxel_grid_splits_each_layer_in_two():
    # a 0.5 mm grid: the beam's steps (chords of the 1 mm cells) are cut into
    # segments no longer than 0.25 mm, so each half-millimetre voxel holds the
    # loss over its own half layer, up to the segments that straddle z = k + 0.5
    om = O.OracleMesh(*beam_box(), nodal=False, isreflect=0)
    cfg = beam_sim(n=16, step=0.5)
    o = O.run(om, cfg, photonseed=seeds_of(cfg.seed, np.arange(cfg.nphoton)), math="portable")
    nmin, dim = O.grid(om, 0.5)
    assert dim == [21, 21, 41]
    f = o.field.reshape(dim[2], dim[1], dim[0])
    col = f[:, 7, 10]
    assert np.count_nonzero(f) == np.count_nonzero(col) == 40
    pair = col[:40].reshape(20, 2).sum(1)
    k = np.arange(20)
    assert np.allclose(pair, cfg.nphoton * (np.exp(-MU * k) - np.exp(-MU * (k + 1))), rtol=2e-5)


def test_grid_only_moves_deposits():
    # same photons, element output vs grid output: identical paths, and the
    # energy deposited sums to the same total
    arrays = meshgen.slab_layers(2.0, nx=40, ny=40, nz=20)
    om = O.OracleMesh(*arrays, nodal=False)
    base = dict(nphoton=3000, seed=21, srcpos=(20.0, 20.0, 0.0), srcdir=(0.0, 0.0, 1.0), tstart=0.0,
                tend=5e-9, tstep=1e-9, outputtype="e")
    ce = M.SimConfig(method="S", **base)
    cg = M.SimConfig(method="G", gridstep=0.7, **base)
    sd = seeds_of(ce.seed, np.arange(ce.nphoton))
    oe = O.run(om, ce, photonseed=sd, math="portable", records=(0, ce.nphoton))
    og = O.run(om, cg, photonseed=sd, math="portable", records=(0, ce.nphoton))
    for key in ("p", "v", "eid", "reason", "nraytet", "w", "t"):
        assert np.array_equal(oe.photons[key], og.photons[key])
    assert og.field.sum() == pytest.approx(oe.field.sum(), rel=1e-6)
    # per gate too: the time gate of a step is the same in both outputs
    assert np.allclose(og.field.sum(1), oe.field.sum(1), rtol=1e-6)


@pytest.mark.gpu
def test_gpu_grid_matches_the_oracle_on_the_beam():
    arrays = beam_box()
    tm = M.TetMesh(*arrays).prepare(nodal=False, isreflect=0)
    om = O.OracleMesh(*arrays, nodal=False, isreflect=0)
    for step in (1.0, 0.5):
        cfg = beam_sim(n=256, step=step)
        # compensated atomics: the FP32 field is then exact to ~1e-9, so the
        # comparison sees the deposits, not the accumulation order
        g = M.run(tm, cfg, records=(0, cfg.nphoton), tune=M.sim.tune_struct(compensated=1))
        o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), math="portable")
        assert g.field.shape == o.field.shape
        gz, oz = np.flatnonzero(g.field), np.flatnonzero(o.field)
        rel = np.abs(g.field - o.field).max() / np.abs(o.field).max()
        print(f"step {step}: nonzero voxels gpu {gz.tolist()[:8]}.. ({len(gz)}), oracle {oz.tolist()[:8]}.. "
              f"({len(oz)}), max rel diff {rel:.3e}", flush=True)
        assert np.array_equal(gz, oz)
        assert np.allclose(g.field, o.field, rtol=1e-6, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("out", ["x", "e"])
def test_gpu_grid_matches_the_oracle_on_the_slab(out):
    arrays = meshgen.slab_layers(2.0)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    om = O.OracleMesh(*arrays, nodal=False)
    cfg = M.SimConfig(nphoton=20000, seed=77, method="G", gridstep=1.0, srcpos=(50.0, 50.0, 0.0),
                      srcdir=(0.0, 0.0, 1.0), tstart=0.0, tend=5e-9, tstep=5e-10, outputtype=out)
    g = M.run(tm, cfg, records=(0, cfg.nphoton), normalize=True)
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              math="portable", normalize=True)
    same = (g.photons["eid"] == o.photons["eid"]) & (g.photons["reason"] == o.photons["reason"])
    assert same.mean() >= 0.999
    rel = np.linalg.norm(g.field - o.field) / np.linalg.norm(o.field)
    print(f"grid field rel L2 {rel:.3e}, normalizer {g.normalizer:.6e} vs {o.normalizer:.6e}", flush=True)
    assert rel < 1e-4
    assert abs(g.normalizer / o.normalizer - 1) < 1e-6
    # and against the reference's arithmetic (the host C library's libm)
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              math="glibc", normalize=True)
    same = (g.photons["eid"] == o.photons["eid"]) & (g.photons["reason"] == o.photons["reason"])
    rel = np.linalg.norm(g.field - o.field) / np.linalg.norm(o.field)
    print(f"against the glibc oracle: same photons {same.mean():.5f}, grid field rel L2 {rel:.3e}", flush=True)
    assert same.mean() >= 0.999 and rel < 1e-4


@pytest.mark.gpu
def test_gpu_grid_batch_equals_runs_alone():
    arrays = meshgen.slab_layers(2.0)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    cfgs = [M.SimConfig(nphoton=5000, seed=5 + i, method="G", gridstep=step, srcpos=(40.0 + 10 * i, 50.0, 0.0),
                        srcdir=(0.0, 0.0, 1.0), tstart=0.0, tend=5e-9, tstep=1e-9, outputtype="x")
            for i, step in enumerate((1.0, 2.5))]
    rm = M.ResidentMesh(tm)
    try:
        batch = rm.run_batch(cfgs)
    finally:
        rm.release()
    for c, b in zip(cfgs, batch):
        a = M.run(tm, c)
        assert b.field.shape == a.field.shape == (5, c.datalen(tm))
        assert np.linalg.norm(b.field - a.field) <= 1e-5 * np.linalg.norm(a.field)
