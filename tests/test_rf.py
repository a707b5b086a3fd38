"""RF forward mode (omega > 0; src/mmc_core.cl, IS_RF, GPU-only in the
reference): the photon weight is complex and turns by omega n / c0 per unit
path while it decays by mua; each step deposits (w_before - w_after) /
(mua + i omega n / c0), real part in the field, imaginary part in field_im.

Pin: along a non-scattering beam the deposits telescope, so voxel layer k of
a 1 mm grid (and the sum over the elements of cell layer k) holds exactly
(e^{-zk} - e^{-z(k+1)}) / z per photon with z = mua + i omega n / c0; and the
magnitude of the weight decays as without modulation, so the energy ledger
(launched = absorbed + escaped) holds with magnitudes.
"""
import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import meshgen
from mmc_amd.replay import seeds_of

MU = 0.1
KAPPA = 0.2                                   # omega * n / c0, 1/mm
R_C0 = 3.335640951981520e-12                  # s/mm, src/mmc_const.h:45
OMEGA = KAPPA / R_C0


def beam_box():
    node, elem = meshgen.t6_lattice(10, 10, 20, 1.0)
    etype = np.ones(len(elem), np.int32)
    med = np.array([[0, 0, 1, 1], [MU, 0.0, 0.0, 1.0]], np.float32)
    return node, elem, etype, med


def beam_sim(method, n=32):
    return M.SimConfig(nphoton=n, seed=3, method=method, gridstep=1.0, srcpos=(5.25, 3.75, 0.0),
                       srcdir=(0.0, 0.0, 1.0), tstart=0.0, tend=1e-8, tstep=1e-8, outputtype="f",
                       isreflect=0, minenergy=0.0, omega=OMEGA)


def layer_sums(arrays, field):
    """Per cell layer k: the sum over the elements whose centroid lies in it."""
    node, elem = arrays[0], arrays[1]
    k = np.floor(meshgen.centroids(node, elem)[:, 2]).astype(int)
    return np.bincount(k, weights=field, minlength=20)


def expected(n):
    z = MU + 1j * KAPPA
    k = np.arange(20)
    return n * (np.exp(-z * k) - np.exp(-z * (k + 1))) / z


@pytest.mark.parametrize("method", ["G", "S"])
def test_beam_complex_fluence_closed_form(method):
    arrays = beam_box()
    om = O.OracleMesh(*arrays, nodal=False, isreflect=0)
    cfg = beam_sim(method)
    o = O.run(om, cfg, photonseed=seeds_of(cfg.seed, np.arange(cfg.nphoton)), math="portable")
    assert o.field_im is not None
    if method == "G":
        nmin, dim = O.grid(om, 1.0)
        col = lambda f: f.reshape(dim[2], dim[1], dim[0])[:20, 3, 5]
        got = col(o.field[0]) + 1j * col(o.field_im[0])
        assert np.count_nonzero(o.field) == 20
    else:
        got = layer_sums(arrays, o.field[0]) + 1j * layer_sums(arrays, o.field_im[0])
    want = expected(cfg.nphoton)
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()
    # the weight's magnitude decays as without modulation
    assert o.absorbweight == pytest.approx(cfg.nphoton * (1 - np.exp(-MU * 20)), rel=1e-5)


def test_rf_energy_ledger_with_scattering():
    arrays = meshgen.slab_layers(2.0, nx=40, ny=40, nz=20)
    om = O.OracleMesh(*arrays, nodal=False)
    cfg = M.SimConfig(nphoton=2000, seed=6, method="G", gridstep=1.0, srcpos=(20.0, 20.0, 0.0),
                      srcdir=(0.0, 0.0, 1.0), tstart=0.0, tend=5e-9, tstep=5e-9, outputtype="f", omega=OMEGA)
    o = O.run(om, cfg, photonseed=seeds_of(cfg.seed, np.arange(cfg.nphoton)), math="portable",
              records=(0, cfg.nphoton))
    ph = o.photons
    led = ph["eabs"].astype(np.float64) + ph["w"] - ph["rgain"]
    assert np.allclose(led, ph["w0"], rtol=0, atol=2e-5)
    assert np.abs(o.field_im).sum() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["G", "S"])
def test_gpu_rf_matches_the_oracle(method):
    arrays = meshgen.slab_layers(2.0)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    om = O.OracleMesh(*arrays, nodal=False)
    cfg = M.SimConfig(nphoton=20000, seed=17, method=method, gridstep=1.0, srcpos=(50.0, 50.0, 0.0),
                      srcdir=(0.0, 0.0, 1.0), tstart=0.0, tend=5e-9, tstep=1e-9, outputtype="f", omega=OMEGA)
    g = M.run(tm, cfg, records=(0, cfg.nphoton), normalize=True, tune=M.sim.tune_struct(compensated=1))
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              math="portable", normalize=True)
    same = (g.photons["eid"] == o.photons["eid"]) & (g.photons["reason"] == o.photons["reason"])
    assert same.mean() >= 0.999
    for a, b in ((g.field, o.field), (g.field_im, o.field_im)):
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"{method}: rel L2 {rel:.3e}", flush=True)
        assert rel < 1e-4
    assert g.absorbweight == pytest.approx(o.absorbweight, rel=1e-5)


@pytest.mark.gpu
def test_gpu_rf_beam_closed_form():
    arrays = beam_box()
    tm = M.TetMesh(*arrays).prepare(nodal=False, isreflect=0)
    cfg = beam_sim("G", n=256)
    g = M.run(tm, cfg, tune=M.sim.tune_struct(compensated=1))
    nmin, dim = cfg.grid(tm)
    col = lambda f: f.reshape(dim[2], dim[1], dim[0])[:20, 3, 5]
    got = col(g.field[0]) + 1j * col(g.field_im[0])
    want = expected(cfg.nphoton)
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()
