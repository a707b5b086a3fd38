"""Per-photon parity with the reference's arithmetic (glibc libm).

The reference CPU tracer (mmc_run_mp, built as its `sse` target without
-ffast-math) computes every scattering length, azimuth and isotropic zenith
with glibc's logf / sincosf / acosf.  The oracle's default build calls glibc
exactly there, and the HIP kernel reproduces glibc's values on every
generator output through its agreement tables.  Both photons of a pair start
from the same xorshift128+ state, so:

* each photon ends in the same element (>= 99.9 %) at the same position
  (|dp| <= 1e-4 mm on >= 99.9 %; in practice bit for bit),
* its final weight agrees to 1e-4 relative (glibc's expf and the kernel's FP32
  expf differ by one ulp on ~0.4 % of arguments; weights do not feed back into
  positions except through the roulette threshold),
* floating-point tie cases -- ray-tetrahedron tests whose exit face a tie rule
  decided, and fixphoton nudges -- are counted by the oracle and reported, and
  photons with ties are held to the same bar.

Workloads: the 60 mm cube of 2 mm cells (Plucker, nodal, roulette), the
full 100x100x50 mm slab of 1 mm cells (3.0M tets; branch-less Badouel,
element output, 16 detectors), and every source type on a scattering cube.
"""
import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import workloads as W
from test_sources import CASES, source_case

pytestmark = pytest.mark.gpu


def report(g, o, ties, tag):
    same = g["eid"] == o["eid"]
    dp = np.abs(g["p"].astype(np.float64) - o["p"]).max(axis=1)
    close = dp <= 1e-4
    wok = np.abs(g["w"].astype(np.float64) - o["w"]) <= 1e-4 * np.maximum(np.abs(o["w"]), 1e-30)
    tied = (ties[:, 0] > 0) | (ties[:, 1] > 0)
    print(f"{tag}: {len(g)} photons; same exit element {same.mean():.5f}, |dp|<=1e-4 mm {close.mean():.5f}, "
          f"bitwise position {(dp == 0).mean():.5f}, weight within 1e-4 {wok.mean():.5f}; "
          f"tie cases: {int(tied.sum())} photons ({int(ties[:, 0].sum())} tie-decided tests, "
          f"{int(ties[:, 1].sum())} nudges), of which {int((tied & ~(same & close)).sum())} differ", flush=True)
    return same, close, wok, tied


@pytest.fixture(scope="module")
def cube():
    arrays = M.meshgen.cube60(2.0)
    return M.TetMesh(*arrays).prepare(nodal=True), O.OracleMesh(*arrays, nodal=True)


@pytest.fixture(scope="module")
def slab_full():
    arrays = M.meshgen.slab_layers(1.0)
    return arrays, M.TetMesh(*arrays).prepare(nodal=False), O.OracleMesh(*arrays, nodal=False)


def test_cube_plucker_nodal_matches_glibc_per_photon(cube):
    tm, om = cube
    cfg = W.cube60_sim(20000)
    seeds = M.photon_seeds(cfg.seed, 0, cfg.nphoton)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=seeds, records=(0, cfg.nphoton), math="glibc")
    same, close, wok, tied = report(g.photons, o.photons, o.ties, "cube 2 mm, Plucker nodal")
    assert same.mean() >= 0.999 and close.mean() >= 0.999
    assert wok.mean() >= 0.999
    assert np.array_equal(g.photons["reason"], o.photons["reason"])
    assert tied.sum() > 0      # the pencil starts on a lattice node: every photon is nudged once
    assert (same & close)[tied].mean() >= 0.999


def test_full_slab_badouel_matches_glibc_per_photon_and_detectors(slab_full):
    arrays, tm, om = slab_full
    assert tm.ne == 3_000_000
    cfg = W.slab_sim(20000, seed=99)
    seeds = M.photon_seeds(cfg.seed, 0, cfg.nphoton)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=seeds, records=(0, cfg.nphoton), math="glibc", maxdet=cfg.nphoton)
    same, close, wok, tied = report(g.photons, o.photons, o.ties, "slab 1 mm (3.0M tets), Badouel element")
    assert same.mean() >= 0.999 and close.mean() >= 0.999
    assert wok.mean() >= 0.999
    assert (same & close)[tied].mean() >= 0.999
    assert np.array_equal(g.photons["nraytet"], o.photons["nraytet"])
    # detected photons: same ids, same detector, scattering counts and partial paths
    assert g.detected == o.detected > 0
    gi, oi = np.argsort(g.detphoton), np.argsort(o.detphoton)
    assert np.array_equal(g.detphoton[gi], o.detphoton[oi])
    gd, od = g.detdata[gi], o.detdata[oi]
    prop = len(arrays[3]) - 1
    assert np.array_equal(gd[:, :1 + prop], od[:, :1 + prop])          # detector id, scattering counts
    assert np.allclose(gd[:, 1 + prop:], od[:, 1 + prop:], rtol=1e-6, atol=0)   # partial paths, w0


@pytest.mark.parametrize("name", sorted(CASES))
def test_every_source_type_matches_glibc_per_photon(name):
    # the launch of each source type (src/mmc_raytrace.c:2346-2592) followed
    # by transport, against the oracle with the host C library's libm
    tm, om, cfg = source_case(name)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              math="glibc")
    same, close, wok, tied = report(g.photons, o.photons, o.ties, f"source {name}")
    # The z-Gaussian zenith sqrt(-2 log u1) (1 - 2 u2) sigma is a function of
    # two draws, so the agreement tables (functions of one generator output)
    # do not cover its sine and cosine: the kernel rounds them correctly
    # through double, where glibc's sincosf misrounds now and then, and 0.2 %
    # of the photons drift apart by more than 1e-4 mm after the first
    # scattering (none leaves through another element).  The reference's own
    # z-Gaussian takes u2 from an unseeded thread-private state, ran0
    # (src/mmc_host.c:227,265,280-283), so it has no reproducible photons to
    # match anyway.  The launch direction, not a tie, decides which photons
    # drift, so the photons with tie cases (a few hundred) hold the same
    # share up to sampling: 6 of 280 differ.
    zg = name == "zgaussian"
    assert same.mean() >= 0.999 and close.mean() >= (0.995 if zg else 0.999)
    assert wok.mean() >= 0.999
    assert (g.photons["reason"] == o.photons["reason"])[same & close].all()
    if tied.any():
        assert (same & close)[tied].mean() >= (0.95 if zg else 0.999)
