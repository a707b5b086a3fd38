"""The benchmark workloads at their stated mesh sizes, against the oracle
built with the host C library's libm (the reference CPU tracer's arithmetic)
and against themselves (batched or served == run alone).

* batch: all 256 simulations of the batched workload on the 1 mm, 3.0M-tet
  slab in one launch at reduced photon counts -- every simulation's energies
  equal its run alone, a sample matches the oracle photon by photon and
  field by field -- and the same 256 on the 2 mm slab with every field
  compared to its run alone;
* serving: requests on each of the three resident meshes (60 mm cube, 1 mm
  slab, 5.2M-tet jittered box) served by the continuous-batching loop,
  each field against the oracle's;
* large: the 24.0M-tet jittered shells mesh with its 64x64 pattern source,
  100 gates, nodal output and 32 detectors, photon by photon against the
  oracle, and 8 pattern requests served == run alone;
* fluence at 1e7 photons on the 1 mm slab against a fixture of the oracle
  (tests/golden/make_slab_fluence.py): the GPU replays the fixture's photon
  streams, so the two fields differ by far less than two independent 1e7
  runs do.
"""
import os

import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import meshgen
from mmc_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def same_photons(g, o, tag):
    same = g["eid"] == o["eid"]
    dp = np.abs(g["p"].astype(np.float64) - o["p"]).max(axis=1)
    ok = same & (dp <= 1e-4) & (g["reason"] == o["reason"])
    print(f"{tag}: {len(g)} photons, same element {same.mean():.5f}, |dp|<=1e-4 {(dp <= 1e-4).mean():.5f}, "
          f"mismatches {int((~ok).sum())}", flush=True)
    return ok.mean()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def with_media(om, med):
    """The oracle mesh with a simulation's own media (same count); the mesh
    owns its media array (the fixtures hand it a copy)."""
    om.med[:len(med)] = med
    return om


@pytest.fixture(scope="module")
def slab1():
    arrays = meshgen.slab_layers(1.0)
    own = arrays[:3] + (arrays[3].copy(),)
    return arrays, M.TetMesh(*arrays).prepare(nodal=False), O.OracleMesh(*own, nodal=False)


@pytest.mark.timeout(600)
def test_batch_of_256_on_the_full_slab(slab1):
    arrays, tm, om = slab1
    assert tm.ne == 3_000_000
    sims = W.slab_batch_sims(256, lo=1000, hi=4000)
    rm = M.ResidentMesh(tm)
    together = rm.run_batch(sims, want_field=False)
    for s, t in zip(sims, together):
        a = rm.run_batch([s], want_field=False)[0]
        assert t.launchweight == a.launchweight == s.nphoton
        assert abs(t.absorbweight - a.absorbweight) <= 1e-6 * a.launchweight
        assert t.raytet == a.raytet
    # a sample against the oracle, photon by photon and field by field
    for k in (0, 101, 255):
        s = sims[k]
        g = rm.run_batch([s], records=(0, s.nphoton))[0]
        o = O.run(with_media(om, s.med), s, photonseed=M.photon_seeds(s.seed, 0, s.nphoton),
                  records=(0, s.nphoton), math="glibc")
        assert same_photons(g.photons, o.photons, f"batch simulation {k}") >= 0.999
        assert rel(g.field, o.field) < 2e-3
        assert abs(g.absorbweight / o.absorbweight - 1) < 1e-3
    rm.release()
    om.med[:] = arrays[3]


@pytest.mark.timeout(600)
def test_batch_of_256_fields_equal_runs_alone():
    tm = M.TetMesh(*meshgen.slab_layers(2.0)).prepare(nodal=False)
    sims = W.slab_batch_sims(256, lo=1000, hi=4000)
    rm = M.ResidentMesh(tm)
    together = rm.run_batch(sims)
    for s, t in zip(sims, together):
        a = rm.run_batch([s])[0]
        assert t.launchweight == a.launchweight
        assert np.allclose(t.field, a.field, rtol=1e-5, atol=1e-9 * np.abs(a.field).max())
    rm.release()


@pytest.mark.timeout(600)
def test_served_requests_on_three_meshes_match_the_oracle():
    from mmc_amd import serve as S
    arrays = {"cube": meshgen.cube60(2.0), "slab": meshgen.slab_layers(1.0),
              "jitter": meshgen.jittered_box(120, 120, 60, 1.0, amp=0.2)}
    meshes = {k: M.TetMesh(*a).prepare(nodal=False) for k, a in arrays.items()}
    assert meshes["jitter"].ne == 5_184_000
    stream = S.poisson_requests(meshes, 40, rate=200.0, photons=(2000, 6000), seed=11)
    reqs = []
    for key in ("cube", "slab", "jitter"):
        reqs += [r for r in stream if r.mesh == key][:2]
    for r in reqs:
        r.want_field = True
    srv = S.Server(meshes, arena_bytes=4 << 30, replicas=2)
    stats = srv.serve(reqs)
    srv.close()
    assert stats["requests"] == 6
    oms = {}
    for r in reqs:
        if r.mesh not in oms:
            a = arrays[r.mesh]
            oms[r.mesh] = O.OracleMesh(a[0], a[1], a[2], a[3].copy(), nodal=False)
        c = r.cfg
        o = O.run(with_media(oms[r.mesh], c.med), c, photonseed=M.photon_seeds(c.seed, 0, c.nphoton), math="glibc")
        print(f"{r.mesh}: {c.nphoton} photons, {c.gates()[0]} gates, field rel {rel(r.field, o.field):.2e}",
              flush=True)
        assert r.launchweight == o.launchweight
        assert abs(r.absorbweight / o.absorbweight - 1) < 1e-3
        assert rel(r.field, o.field) < 2e-3


@pytest.fixture(scope="module")
def shells():
    node, elem, et, med, srcelem = W.shells_arrays(1.0)
    tm = M.TetMesh(node, elem, et, med, srcelem=srcelem).prepare(nodal=True)
    om = O.OracleMesh(node, elem, et, med, nodal=True, srcelem=srcelem)
    return tm, om


@pytest.mark.timeout(900)
def test_large_shells_mesh_per_photon(shells):
    tm, om = shells
    assert tm.ne == 24_000_000
    cfg = W.shells_sims(8, photons=80000)[0]
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              maxdet=cfg.nphoton, math="glibc")
    assert same_photons(g.photons, o.photons, "shells 24.0M tets, pattern, Plucker nodal") >= 0.999
    assert g.field.shape == o.field.shape == (100, tm.nn)
    assert rel(g.field, o.field) < 2e-3
    assert abs(g.launchweight / o.launchweight - 1) < 1e-6
    assert o.detected > 0 and abs(g.detected - o.detected) <= 2


@pytest.mark.timeout(900)
def test_large_shells_pattern_requests_served_equal_alone(shells):
    from mmc_amd import serve as S
    tm, _ = shells
    sims = W.shells_sims(8, photons=80000)
    reqs = [S.Request(rid=i, mesh="shells", cfg=c, arrival=0.0, want_field=True) for i, c in enumerate(sims)]
    srv = S.Server({"shells": tm}, arena_bytes=16 << 30, replicas=2)
    stats = srv.serve(reqs)
    srv.close()
    assert stats["requests"] == 8
    for r in reqs:
        a = M.run(tm, r.cfg)
        assert r.launchweight == a.launchweight
        assert abs(r.absorbweight - a.absorbweight) <= 1e-5 * a.launchweight
        assert rel(r.field, a.field) < 1e-4
        r.field = None


@pytest.mark.timeout(600)
def test_slab_fluence_at_1e7_against_the_oracle_fixture(slab1):
    arrays, tm, _ = slab1
    fx = np.load(os.path.join(GOLDEN, "slab_fluence_1e7.npz"))
    n = int(fx["photons"])
    assert n >= 10_000_000
    cfg = W.slab_sim(n, seed=int(fx["seed_a"]))
    cfg.issavedet, cfg.detpos = 0, None
    g = M.run(tm, cfg)
    c = M.meshgen.centroids(arrays[0], arrays[1])
    b = np.floor(c / float(fx["block"])).astype(np.int64)
    bidx = b[:, 0] + 20 * (b[:, 1] + 20 * b[:, 2])
    gb = np.stack([np.bincount(bidx, weights=f, minlength=4000) for f in g.field])
    fa, fb = fx["field_a"], fx["field_b"]
    spread = rel(fb, fa)                  # two independent 1e7-photon oracle runs
    d = rel(gb, fa)                       # the GPU on run a's photon streams
    print(f"1e7 photons, 5 mm blocks x 50 gates: GPU vs oracle {d:.2e}, oracle vs oracle (other seed) "
          f"{spread:.2e}", flush=True)
    assert d < 0.1 * spread
    assert abs(g.absorbweight / float(fx["absorb_a"]) - 1) < 1e-4
