"""Physical invariants of the GPU transport that hold whatever the oracle does.

* Mirror symmetry: a pencil beam entering the centre of the homogeneous
  60 mm cube (Fang 2010's validation case, examples/validation) along +z
  deposits, on average, the same energy on both sides of the planes x = 30
  and y = 30.  Element deposits add up to the energy absorbed in a region
  exactly (no interpolation), and these planes are lattice planes, so the
  half-space sums do not depend on how the cells are cut into tetrahedra.
  The pencil starts on the lattice line x = y = 30 and flies along it, on
  faces shared by elements of both sides, so what it absorbs before its
  first scattering goes to the elements of the side the first-match rule
  picked (1.9 % of the absorbed energy here, all on one side): the column of
  cells around that line is left out of both sums.  The asymmetry of eight
  independent runs is tested against its own scatter (a t test).
* Time of flight: a detected photon's time equals the sum over media of its
  partial path times n / c0 (the partial paths of its detection record).
* Exit state: a detected photon's saved exit position lies on the detector
  surface inside its detector, and its direction is a unit vector pointing
  out of the mesh.
"""
import numpy as np
import pytest

import mmc_amd as M
from mmc_amd import meshgen
from mmc_amd import workloads as W

pytestmark = pytest.mark.gpu

R_C0 = 3.335640951981520e-12          # s/mm, src/mmc_const.h:45


def test_absorbed_energy_is_mirror_symmetric():
    node, elem, et, med = meshgen.cube60(2.0)
    tm = M.TetMesh(node, elem, et, med).prepare(nodal=False)
    c = meshgen.centroids(node, elem)
    away = (np.abs(c[:, 0] - 30.0) > 2.0) | (np.abs(c[:, 1] - 30.0) > 2.0)
    halves = {"x": (away & (c[:, 0] < 30.0), away & (c[:, 0] > 30.0)),
              "y": (away & (c[:, 1] < 30.0), away & (c[:, 1] > 30.0))}
    asym = {k: [] for k in halves}
    for i in range(8):
        cfg = M.SimConfig(nphoton=500_000, seed=1000 + i, method="S", srcpos=(30.0, 30.0, 0.0),
                          srcdir=(0.0, 0.0, 1.0), tstart=0.0, tend=5e-9, tstep=5e-9, outputtype="e")
        e = M.run(tm, cfg).field.sum(0)
        for k, (lo, hi) in halves.items():
            a, b = e[lo].sum(), e[hi].sum()
            asym[k].append((a - b) / (a + b))
    for k, v in asym.items():
        v = np.array(v)
        t = v.mean() / (v.std(ddof=1) / np.sqrt(len(v)))
        print(f"{k}: asymmetry {v.mean():+.2e} +- {v.std(ddof=1) / np.sqrt(len(v)):.1e} (t = {t:+.2f})", flush=True)
        assert abs(t) < 6.0                          # P(|t_7| > 6) < 1e-3
        assert abs(v.mean()) < 5e-3                  # about 7 standard errors


def test_detected_time_of_flight_is_path_times_index():
    arrays = meshgen.slab_layers(2.0)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    med = arrays[3]
    prop = len(med) - 1
    cfg = W.slab_sim(20000, seed=61)
    cfg.tstep = cfg.tend                             # one gate: every photon runs to the end of the window
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    assert g.detected > 50
    ppath = g.detdata[:, 1 + prop:1 + 2 * prop].astype(np.float64)
    t_path = (ppath * med[1:, 3].astype(np.float64)).sum(1) * R_C0
    t_rec = g.photons["t"][g.detphoton].astype(np.float64)
    rel = np.abs(t_rec - t_path) / t_path
    print(f"{g.detected} detected photons: time vs path x n / c0 max rel {rel.max():.2e}", flush=True)
    assert rel.max() < 1e-4


def test_exit_positions_and_directions_of_detected_photons():
    # examples/regression/exitangle: with the exit state saved (-x 1), a
    # detected photon left through the surface z = 0 that holds the
    # detectors, inside its detector's disk, moving outward with a unit
    # direction
    arrays = meshgen.slab_layers(2.0)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    prop = len(arrays[3]) - 1
    cfg = W.slab_sim(20000, seed=62)
    cfg.issaveexit = 1
    g = M.run(tm, cfg)
    assert g.detected > 50
    d = g.detdata.astype(np.float64)
    p, v = d[:, 1 + 2 * prop:4 + 2 * prop], d[:, 4 + 2 * prop:7 + 2 * prop]
    det = np.asarray(cfg.detpos, np.float64)[d[:, 0].astype(int) - 1]
    print(f"{g.detected} detected: |z| max {np.abs(p[:, 2]).max():.1e}, vz max {v[:, 2].max():.3f}, "
          f"| |v| - 1 | max {np.abs(np.linalg.norm(v, axis=1) - 1).max():.1e}", flush=True)
    assert np.abs(p[:, 2]).max() < 1e-4
    assert (v[:, 2] < 0).all()
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-5)
    assert (((p[:, :2] - det[:, :2]) ** 2).sum(1) < det[:, 3] ** 2 + 1e-4).all()
