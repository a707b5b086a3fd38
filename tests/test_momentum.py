"""Momentum transfer of detected photons (-m / Session.DoDCS; src/mmc_mesh.c:
1707-1709, src/mmc_raytrace.c:2073-2075): each scattering event in medium k
adds 1 - cos(theta) to the record's third block of prop values, after the
scattering counts and partial paths.

Pin: for Henyey-Greenstein scattering E[cos theta] = g, so over many
detected photons the momentum block divided by the scattering counts tends
to 1 - g in each medium; every entry lies in [0, 2 x count].  The GPU
records equal the oracle's photon by photon.
"""
import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import meshgen


def two_media():
    node, elem = meshgen.t6_lattice(15, 15, 15, 4.0)
    c = meshgen.centroids(node, elem)
    etype = np.where(c[:, 2] < 12.0, 1, 2).astype(np.int32)
    med = np.array([[0, 0, 1, 1], [0.01, 1.0, 0.8, 1.37], [0.005, 0.5, 0.3, 1.37]], np.float32)
    return node, elem, etype, med


def cfg_of(n, seed=8):
    det = np.array([[30.0, 22.0, 0.0, 3.0], [30.0, 38.0, 0.0, 3.0], [22.0, 30.0, 0.0, 3.0]], np.float32)
    return M.SimConfig(nphoton=n, seed=seed, method="S", srcpos=(30.2, 30.1, 0.0), srcdir=(0.0, 0.0, 1.0),
                       tend=5e-9, tstep=5e-9, issavedet=1, detpos=det, ismomentum=1)


def test_oracle_momentum_mean_is_one_minus_g():
    node, elem, etype, med = two_media()
    om = O.OracleMesh(node, elem, etype, med, nodal=False)
    cfg = cfg_of(12000)
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), maxdet=cfg.nphoton, math="portable")
    d = o.detdata.astype(np.float64)
    prop = 2
    assert d.shape[1] == 1 + 3 * prop + 1 and len(d) > 100
    nscat, mom = d[:, 1:1 + prop], d[:, 1 + 2 * prop:1 + 3 * prop]
    assert np.all(mom >= 0) and np.all(mom <= 2 * nscat + 1e-3)
    for k, g in enumerate(med[1:, 2]):
        ratio = mom[:, k].sum() / nscat[:, k].sum()
        print(f"medium {k + 1}: g {g}, momentum per scattering {ratio:.4f}, events {nscat[:, k].sum():.0f}")
        # standard error of the mean of 1 - cos(theta): var(cos) = (1 + 2 g^2) / 3 - g^2 for HG
        se = np.sqrt(((1 + 2 * g * g) / 3 - g * g) / nscat[:, k].sum())
        assert ratio == pytest.approx(1 - g, abs=4 * se + 0.005)


@pytest.mark.gpu
def test_gpu_momentum_records_match_the_oracle():
    node, elem, etype, med = two_media()
    tm = M.TetMesh(node, elem, etype, med).prepare(nodal=False)
    om = O.OracleMesh(node, elem, etype, med, nodal=False)
    cfg = cfg_of(20000, seed=31)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              maxdet=cfg.nphoton, math="portable")
    assert g.detected == o.detected > 0
    assert g.detdata.shape[1] == o.detdata.shape[1] == 1 + 3 * 2 + 1
    gi, oi = np.argsort(g.detphoton), np.argsort(o.detphoton)
    assert np.array_equal(g.detphoton[gi], o.detphoton[oi])
    assert np.allclose(g.detdata[gi], o.detdata[oi], rtol=1e-5, atol=1e-5)
    assert g.detdata[:, 5:7].sum() > 0
