"""GPU parity: the HIP transport kernel against the CPU oracle.

Per-photon checks use the oracle built with the same deterministic FP32 libm
(include/mmc_fp32math.h): every photon must end at the same element, at the
same position within 1e-4 mm, for the same reason.  Fluence checks against
the glibc-libm oracle (the reference's libm) are statistical: the relative L2
difference GPU-vs-CPU must be within 1.5x the spread of two independent-seed
CPU runs.
"""
import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import workloads as W

pytestmark = pytest.mark.gpu


def seeds(cfg):
    return M.photon_seeds(cfg.seed, cfg.photon_offset, cfg.nphoton)


def compare_photons(g, o):
    same_eid = g["eid"] == o["eid"]
    dp = np.abs(g["p"].astype(np.float64) - o["p"]).max(axis=1)
    ok = same_eid & (dp <= 1e-4) & (g["reason"] == o["reason"])
    print(f"photons {len(g)}: same element {same_eid.mean():.5f}, |dp|<=1e-4 {(dp <= 1e-4).mean():.5f}, "
          f"bitwise position {(dp == 0).mean():.5f}, mismatches {int((~ok).sum())}", flush=True)
    return ok


@pytest.fixture(scope="module")
def cube():
    arrays = M.meshgen.cube60(2.0)
    return arrays, M.TetMesh(*arrays).prepare(nodal=True), O.OracleMesh(*arrays, nodal=True)


@pytest.fixture(scope="module")
def slab():
    arrays = M.meshgen.slab_layers(2.0)
    return arrays, M.TetMesh(*arrays).prepare(nodal=False), O.OracleMesh(*arrays, nodal=False)


def test_cube_plucker_nodal_per_photon(cube):
    _, tm, om = cube
    cfg = W.cube60_sim(20000)
    g = M.run(tm, cfg, records=(0, cfg.nphoton), normalize=True)
    o = O.run(om, cfg, photonseed=seeds(cfg), records=(0, cfg.nphoton), normalize=True, math="portable")
    ok = compare_photons(g.photons, o.photons)
    assert ok.mean() >= 0.999
    assert np.array_equal(g.photons["nraytet"], o.photons["nraytet"])
    assert g.raytet == o.raytet
    rel = np.linalg.norm(g.field - o.field) / np.linalg.norm(o.field)
    assert rel < 1e-3          # FP32 atomics vs FP64 accumulation
    assert abs(g.normalizer / o.normalizer - 1) < 1e-4


@pytest.mark.parametrize("method", ["P", "S"])
def test_nodal_output_across_refracting_layers(method):
    # nodal output (Plucker, and Badouel with its exit-face nodes) in a cube
    # whose upper half has another refractive index: at the internal face
    # photons either reflect (and stay in their element) or refract into the
    # next, and the nodal sums follow them only in the second case
    node, elem = M.meshgen.t6_lattice(15, 15, 15, 2.0)
    c = M.meshgen.centroids(node, elem)
    et = np.where(c[:, 2] < 14.0, 1, 2).astype(np.int32)
    med = np.array([[0, 0, 1, 1], [0.01, 1.0, 0.5, 1.37], [0.02, 2.0, 0.7, 1.6]], np.float32)
    tm = M.TetMesh(node, elem, et, med).prepare(nodal=True)
    om = O.OracleMesh(node, elem, et, med, nodal=True)
    cfg = M.SimConfig(nphoton=20000, method=method, basisorder=1, srcpos=(15.1, 14.9, 0.0), srcdir=(0, 0, 1),
                      tend=5e-9, tstep=1e-9, seed=61)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=seeds(cfg), records=(0, cfg.nphoton), math="portable")
    assert compare_photons(g.photons, o.photons).mean() >= 0.999
    assert g.field.shape == o.field.shape == (5, om.nn)
    rel = np.linalg.norm(g.field - o.field) / np.linalg.norm(o.field)
    print(f"nodal field vs oracle: {rel:.2e}", flush=True)
    assert rel < 1e-4


def test_slab_badouel_element_per_photon_and_detectors(slab):
    _, tm, om = slab
    cfg = W.slab_sim(20000, seed=99)
    g = M.run(tm, cfg, records=(0, cfg.nphoton), normalize=True)
    o = O.run(om, cfg, photonseed=seeds(cfg), records=(0, cfg.nphoton), normalize=True, math="portable",
              maxdet=cfg.nphoton)
    ok = compare_photons(g.photons, o.photons)
    assert ok.mean() >= 0.999
    assert g.detected == o.detected > 0
    gi, oi = np.argsort(g.detphoton), np.argsort(o.detphoton)
    assert np.array_equal(g.detphoton[gi], o.detphoton[oi])
    # detector id, scattering counts, partial paths and launch weight per record
    assert np.allclose(g.detdata[gi], o.detdata[oi], rtol=1e-5, atol=1e-5)
    rel = np.linalg.norm(g.field - o.field) / np.linalg.norm(o.field)
    assert rel < 1e-4


def test_energy_ledger(slab):
    _, tm, _ = slab
    cfg = W.slab_sim(50000, seed=5)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    ph = g.photons
    led = ph["w0"].astype(np.float64) - ph["eabs"] - ph["w"] + ph["rgain"]
    assert np.abs(led).max() < 1e-5
    total = g.absorbweight + g.escapeweight - g.roulettegain
    assert abs(total - g.launchweight) <= 1e-5 * g.launchweight
    # the field holds the same energy: sum of (w0 - w1)/mua deposits times mua
    assert g.fieldsum > 0


def test_detected_weight_follows_beer_lambert(slab):
    arrays, tm, _ = slab
    cfg = W.slab_sim(50000, seed=11)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    med = arrays[3].astype(np.float64)
    prop = len(med) - 1
    ppath = g.detdata[:, 1 + prop:1 + 2 * prop].astype(np.float64)
    w_mbll = g.detdata[:, -1] * np.exp(-(ppath * med[1:, 0]).sum(axis=1))
    w_end = g.photons["w"][g.detphoton].astype(np.float64)
    nstep = g.photons["nraytet"][g.detphoton].astype(np.float64)
    # The kernel multiplies n FP32 exp factors into w and sums the partial
    # paths in FP32, so |ln(w_end/w_mbll)| grows with the step count n and the
    # optical depth E: at most ~n*2^-24*(4 + |E|) (rounding of n products, the
    # <= 1.5 ulp exp, and of n path additions scaled by E).  A dropped or
    # doubled term (a tissue's ppath, one step's attenuation) is orders above.
    E = np.abs(np.log(w_mbll / g.detdata[:, -1]))
    err = np.abs(np.log(w_end / w_mbll))
    assert np.all(err <= 4 * nstep * 2.0 ** -24 * (4 + E)), (err.max(), nstep[err.argmax()], E[err.argmax()])
    assert np.median(err) < 2e-5
    # time of flight = sum_i n_i L_i / c
    t = (ppath * med[1:, 3]).sum(axis=1) * 3.335640951981520e-12
    assert np.allclose(t, g.photons["t"][g.detphoton], rtol=2e-4)


def test_fluence_within_monte_carlo_noise_of_reference_libm(cube):
    arrays, _, _ = cube
    a4 = M.meshgen.cube60(4.0)
    tm = M.TetMesh(*a4).prepare(nodal=True)
    om = O.OracleMesh(*a4, nodal=True)
    cfg = W.cube60_sim(100000)
    g = M.run(tm, cfg, normalize=True)
    c1 = O.run(om, M.SimConfig(**{**cfg.__dict__, "seed": 11}), normalize=True)
    c2 = O.run(om, M.SimConfig(**{**cfg.__dict__, "seed": 12}), normalize=True)
    spread = np.linalg.norm(c1.field - c2.field) / np.linalg.norm(c1.field)
    diff = np.linalg.norm(g.field - c1.field) / np.linalg.norm(c1.field)
    print(f"relative L2: GPU vs CPU {diff:.4g}, CPU vs CPU {spread:.4g}")
    assert diff <= 1.5 * spread
    assert abs(g.absorbweight / g.launchweight - c1.absorbweight / c1.launchweight) < 0.01


def test_roulette_per_photon():
    node, elem, et, med = M.meshgen.cube60(4.0)
    med = med.copy()
    med[1] = [0.3, 2.0, 0.5, 1.37]      # strong absorption: weights fall below 1e-6
    tm = M.TetMesh(node, elem, et, med).prepare(nodal=True)
    om = O.OracleMesh(node, elem, et, med, nodal=True)
    cfg = M.SimConfig(nphoton=20000, method="P", srcpos=(30.1, 30.2, 0), srcdir=(0, 0, 1), tend=5e-9, tstep=5e-9,
                      seed=4)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=seeds(cfg), records=(0, cfg.nphoton), math="portable")
    assert (o.photons["reason"] == 3).sum() > 100
    assert compare_photons(g.photons, o.photons).mean() >= 0.999
    assert abs(g.roulettegain - o.photons["rgain"].astype(np.float64).sum()) < 1e-3 * cfg.nphoton


@pytest.mark.parametrize("srctype", ["planar", "pattern"])
def test_area_sources_per_photon(srctype):
    node, elem, et, med = M.meshgen.cube60(4.0)
    base = M.TetMesh(node, elem, et, med).prepare(nodal=False)
    c = M.meshgen.centroids(base.node, base.elem)
    srcelem = (np.where(c[:, 2] < 4.0)[0] + 1).astype(np.int32)
    tm = M.TetMesh(node, elem, et, med, srcelem=srcelem).prepare(nodal=False)
    om = O.OracleMesh(node, elem, et, med, nodal=False, srcelem=srcelem)
    pat = np.random.default_rng(3).integers(0, 2, size=(8, 8)).astype(np.float32)
    kw = dict(srctype=srctype, srcpos=(20.0, 20.0, 0.0), srcparam1=(20.0, 0, 0, 8), srcparam2=(0, 20.0, 0, 8))
    if srctype == "pattern":
        kw["srcpattern"] = pat
    cfg = M.SimConfig(nphoton=10000, method="S", srcdir=(0, 0, 1), tend=5e-9, tstep=5e-10, seed=21,
                      e0=int(srcelem[0]), **kw)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=seeds(cfg), records=(0, cfg.nphoton), math="portable")
    assert compare_photons(g.photons, o.photons).mean() >= 0.999
    assert abs(g.launchweight - o.launchweight) < 1e-3


@pytest.mark.parametrize("method", ["S", "P"])
def test_shared_patterns_equal_one_run_per_pattern(method):
    # photon sharing (srcnum > 1, src/mmc_raytrace.c:1586-1612, 1812-1835):
    # slot k of a run whose photons carry all patterns equals the run with
    # pattern k alone.  With 0/1 patterns and several gates (no roulette) the
    # single-pattern photons have weight p_k and the same trajectories, so
    # the deposits agree value by value up to the FP32 atomic order.
    node, elem, et, med = M.meshgen.cube60(4.0)
    nodal = method == "P"
    base = M.TetMesh(node, elem, et, med).prepare(nodal=nodal)
    c = M.meshgen.centroids(base.node, base.elem)
    srcelem = (np.where(c[:, 2] < 4.0)[0] + 1).astype(np.int32)
    tm = M.TetMesh(node, elem, et, med, srcelem=srcelem).prepare(nodal=nodal)
    pats = (np.random.default_rng(5).random((3, 8, 8)) < 0.5).astype(np.float32)
    kw = dict(nphoton=20000, method=method, srctype="pattern", srcpos=(20.0, 20.0, 0.0), srcdir=(0, 0, 1),
              srcparam1=(20.0, 0, 0, 8), srcparam2=(0, 20.0, 0, 8), tend=5e-9, tstep=5e-10, seed=31,
              e0=int(srcelem[0]))
    tune = M.sim.tune_struct(compensated=1)
    shared = M.run(tm, M.SimConfig(srcpattern=pats, **kw), tune=tune)
    assert shared.field.shape[-1] == 3 and shared.launchweight == kw["nphoton"]
    for k in range(3):
        one = M.run(tm, M.SimConfig(srcpattern=pats[k], **kw), tune=tune)
        f = shared.field[..., k]
        assert np.linalg.norm(f - one.field) <= 1e-6 * np.linalg.norm(one.field)
        assert shared.srcweight[0, k] == one.launchweight
        assert abs(shared.srcweight[1, k] - one.absorbweight) <= 1e-6 * one.launchweight
    # per-pattern normalization uses each pattern's energies
    k_shared = M.sim.normalize_field(tm, M.SimConfig(srcpattern=pats, **kw), shared)
    one = M.run(tm, M.SimConfig(srcpattern=pats[2], **kw), tune=tune, normalize=True)
    assert abs(k_shared[2] / one.normalizer - 1) < 1e-5
    assert np.linalg.norm(shared.field[..., 2] - one.field) <= 1e-5 * np.linalg.norm(one.field)


@pytest.mark.parametrize("method,basis", [("S", 0), ("S", 1), ("P", 1)])
def test_shared_patterns_match_oracle(method, basis):
    # photon sharing against the oracle's own restatement: same photons, and
    # every pattern's slot of the field and its launched/absorbed energies
    # (element and exit-face-node output under Badouel, nodal output under
    # Plucker)
    nodal = basis == 1
    node, elem, et, med = M.meshgen.cube60(4.0)
    base = M.TetMesh(node, elem, et, med).prepare(nodal=nodal)
    c = M.meshgen.centroids(base.node, base.elem)
    srcelem = (np.where(c[:, 2] < 4.0)[0] + 1).astype(np.int32)
    tm = M.TetMesh(node, elem, et, med, srcelem=srcelem).prepare(nodal=nodal)
    om = O.OracleMesh(node, elem, et, med, nodal=nodal, srcelem=srcelem)
    pats = np.random.default_rng(6).uniform(0.1, 1.0, (4, 6, 8)).astype(np.float32)
    cfg = M.SimConfig(nphoton=20000, method=method, srctype="pattern", srcpos=(20.0, 20.0, 0.0), srcdir=(0, 0, 1),
                      srcparam1=(20.0, 0, 0, 8), srcparam2=(0, 20.0, 0, 6), srcpattern=pats, tend=5e-9,
                      tstep=1e-9, seed=41, e0=int(srcelem[0]), basisorder=basis)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=seeds(cfg), records=(0, cfg.nphoton), math="portable")
    assert compare_photons(g.photons, o.photons).mean() >= 0.999
    assert g.field.shape == o.field.shape == (5, om.nn if nodal else om.ne, 4)
    for k in range(4):
        rel = np.linalg.norm(g.field[..., k] - o.field[..., k]) / np.linalg.norm(o.field[..., k])
        assert rel < 1e-3, (k, rel)
    assert np.allclose(g.srcweight, o.srcweight, rtol=1e-3, atol=0)
    # and against the reference's arithmetic (the host C library's libm)
    o = O.run(om, cfg, photonseed=seeds(cfg), records=(0, cfg.nphoton), math="glibc")
    assert compare_photons(g.photons, o.photons).mean() >= 0.999
    for k in range(4):
        rel = np.linalg.norm(g.field[..., k] - o.field[..., k]) / np.linalg.norm(o.field[..., k])
        assert rel < 1e-3, (k, rel)


def test_batch_result_equals_result_alone(slab):
    _, tm, _ = slab
    sims = W.slab_batch_sims(3, seed=1, lo=2000, hi=6000)
    rm = M.ResidentMesh(tm)
    together = rm.run_batch(sims, records=(0, sims[0].nphoton))
    alone = [rm.run_batch([s], records=(0, s.nphoton) if i == 0 else None)[0] for i, s in enumerate(sims)]
    rm.release()
    assert np.array_equal(together[0].photons, alone[0].photons)
    for a, b in zip(together, alone):
        assert a.launchweight == b.launchweight
        assert abs(a.absorbweight - b.absorbweight) < 1e-6 * a.launchweight
        assert np.allclose(a.field, b.field, rtol=1e-5, atol=1e-9 * np.abs(b.field).max())


def test_zero_and_single_photon(slab):
    _, tm, _ = slab
    r0 = M.run(tm, W.slab_sim(0))
    assert r0.launchweight == 0 and r0.fieldsum == 0
    r1 = M.run(tm, W.slab_sim(1), records=(0, 1))
    assert r1.launchweight == 1.0 and r1.photons["reason"][0] in (1, 2)


def test_compensated_atomics_track_fp64_accumulation(cube):
    _, tm, om = cube
    cfg = W.cube60_sim(50000)
    o = O.run(om, cfg, photonseed=seeds(cfg), math="portable")
    plain = M.run(tm, cfg)
    comp = M.run(tm, cfg, tune=M.sim.tune_struct(compensated=1))
    e_plain = np.linalg.norm(plain.field - o.field) / np.linalg.norm(o.field)
    e_comp = np.linalg.norm(comp.field - o.field) / np.linalg.norm(o.field)
    print(f"FP32 atomics vs FP64: {e_plain:.3g}; compensated: {e_comp:.3g}")
    assert e_comp <= e_plain
    assert e_comp < 1e-6


@pytest.mark.parametrize("replicas", [1, 3, 8])
def test_field_copies_sum_to_the_fp64_accumulation(cube, replicas):
    # the per-XCD field copies only split the atomics; their FP64 sum is the
    # field (compensated copies: every photon's deposit is in exactly one copy)
    _, tm, om = cube
    cfg = W.cube60_sim(30000)
    o = O.run(om, cfg, photonseed=seeds(cfg), math="portable")
    g = M.run(tm, cfg, tune=M.sim.tune_struct(compensated=1, replicas=replicas))
    assert g.launchweight == o.launchweight
    assert np.linalg.norm(g.field - o.field) / np.linalg.norm(o.field) < 1e-6
    assert abs(g.fieldsum - o.field.sum()) <= 1e-6 * abs(o.field.sum())


@pytest.mark.parametrize("policy", ["fair", "srpt"])
def test_served_results_equal_results_alone(policy):
    from mmc_amd.serve import Request, Server
    arrays = M.meshgen.slab_layers(4.0)
    mesh = M.TetMesh(*arrays).prepare(nodal=False)
    cube = M.TetMesh(*M.meshgen.cube60(4.0)).prepare(nodal=False)
    meshes = {"slab": mesh, "cube": cube}
    cfgs = [("slab", W.slab_batch_sims(3, seed=2, lo=30000, hi=120000)[i]) for i in range(3)]
    cfgs.append(("cube", M.SimConfig(nphoton=70000, method="S", srcpos=(30.1, 30.2, 0), srcdir=(0, 0, 1),
                                     tend=5e-9, tstep=1e-9, seed=77)))
    reqs = [Request(rid=i, mesh=k, cfg=c, arrival=0.001 * i, want_field=True) for i, (k, c) in enumerate(cfgs)]
    # compensated (TwoSum) accumulation on both sides, so that the fields can be
    # compared far below the FP32 atomic-ordering noise (~1e-4 relative)
    srv = Server(meshes, arena_bytes=1 << 28, slice_photons=60000, compensated=True, policy=policy)
    srv.sched.min_chunk = 20000          # several chunks per request
    stats = srv.serve(reqs)
    srv.close()
    assert stats["requests"] == len(reqs) and stats["launches"] > len(reqs)
    assert sum(r.launches > 1 for r in reqs) >= 2      # requests were split across launches
    for r in reqs:
        alone = M.run(meshes[r.mesh], r.cfg, tune=M.sim.tune_struct(compensated=1))
        assert r.launchweight == alone.launchweight
        assert abs(r.absorbweight - alone.absorbweight) <= 1e-6 * alone.launchweight
        assert abs(r.raytet - alone.raytet) == 0
        rel = np.linalg.norm(r.field - alone.field) / np.linalg.norm(alone.field)
        assert rel < 1e-6, rel


def test_replay_of_saved_seeds_reproduces_the_detected_photons(slab):
    # seed replay (src/mmc_mesh.c:815-898): relaunching the detected photons
    # from their initial states gives the same detection records
    from mmc_amd import replay as R
    _, tm, _ = slab
    cfg = W.slab_sim(30000, seed=23)
    g = M.run(tm, cfg)
    assert g.detected > 100
    seeds = R.detected_seeds(cfg, g)
    order = np.argsort(g.detphoton)
    rcfg = M.SimConfig(**{**cfg.__dict__, "nphoton": len(seeds), "photonseed": seeds[order]})
    r = M.run(tm, rcfg, records=(0, len(seeds)))
    assert r.detected == len(seeds)
    ri = np.argsort(r.detphoton)
    assert np.array_equal(r.detdata[ri], g.detdata[order])


@pytest.mark.parametrize("out", ["j", "l", "p"])
def test_replay_outputs_match_the_oracle(slab, out):
    # Jacobian, weighted path and weighted scattering count of the replayed
    # detected photons against the oracle's restatement of the same formulas
    from mmc_amd import replay as R
    arrays, tm, om = slab
    cfg = W.slab_sim(30000, seed=29)
    g = M.run(tm, cfg)
    seeds = R.detected_seeds(cfg, g)
    mch = R.MchFile(len(arrays[3]) - 1, len(cfg.detpos), g.detdata.shape[1], cfg.nphoton, g.detected, 1.0, 1.0,
                    g.detdata, seeds)
    rcfg = R.replay_config(cfg, mch, arrays[3], out)
    rcfg.issavedet = 0
    gr = M.run(tm, rcfg, tune=M.sim.tune_struct(compensated=1), normalize=True)
    o = O.run(om, rcfg, photonseed=rcfg.photonseed, math="portable", normalize=True)
    assert gr.launchweight == pytest.approx(o.launchweight, rel=1e-6)
    assert np.linalg.norm(gr.field - o.field) <= 1e-5 * np.linalg.norm(o.field)
    assert abs(gr.normalizer / o.normalizer - 1) < 1e-6
    # and against the reference's arithmetic (the host C library's libm)
    o = O.run(om, rcfg, photonseed=rcfg.photonseed, math="glibc", normalize=True)
    rel = np.linalg.norm(gr.field - o.field) / np.linalg.norm(o.field)
    print(f"replay {out} against the glibc oracle: field rel L2 {rel:.3e}", flush=True)
    assert gr.launchweight == pytest.approx(o.launchweight, rel=1e-6) and rel <= 1e-4


@pytest.mark.parametrize("out", ["l", "p"])
def test_pattern_replay_outputs_match_the_oracle(out):
    # weighted path / scattering outputs of a pattern source: each replayed
    # photon's weight takes its launch pixel's pattern value
    # (src/mmc_raytrace.c:2394-2399), on the GPU as in the oracle
    from mmc_amd import replay as R
    arrays = M.meshgen.slab_layers(4.0)
    c = M.meshgen.centroids(arrays[0], arrays[1])
    srcelem = (np.where(c[:, 2] < 8.0)[0] + 1).astype(np.int32)
    tm = M.TetMesh(*arrays, srcelem=srcelem).prepare(nodal=False)
    om = O.OracleMesh(*arrays, nodal=False, srcelem=srcelem)
    cfg = W.slab_sim(30000, seed=41)
    cfg.srctype, cfg.srcpos = "pattern", (30.3, 30.7, 0.0)
    cfg.srcparam1, cfg.srcparam2 = (40.0, 0.0, 0.0, 8), (0.0, 40.0, 0.0, 8)
    cfg.srcpattern = np.random.default_rng(8).uniform(0.2, 1.0, (8, 8)).astype(np.float32)
    g = M.run(tm, cfg)
    assert g.detected > 100
    seeds_ = R.detected_seeds(cfg, g)
    mch = R.MchFile(len(arrays[3]) - 1, len(cfg.detpos), g.detdata.shape[1], cfg.nphoton, g.detected, 1.0, 1.0,
                    g.detdata, seeds_)
    rcfg = R.replay_config(cfg, mch, arrays[3], out)
    rcfg.issavedet = 0
    gr = M.run(tm, rcfg, tune=M.sim.tune_struct(compensated=1), normalize=True)
    o = O.run(om, rcfg, photonseed=rcfg.photonseed, math="portable", normalize=True)
    # launched energy: the detected weights times their pixels' pattern values
    w0 = g.detdata[:, -1].astype(np.float64)
    assert gr.launchweight == pytest.approx(float((rcfg.replayweight.astype(np.float64) * w0).sum()), rel=1e-5)
    assert gr.launchweight == pytest.approx(o.launchweight, rel=1e-6)
    assert np.linalg.norm(gr.field - o.field) <= 1e-5 * np.linalg.norm(o.field)


def test_forty_media_labels_match_the_oracle_per_photon():
    # more labels than the 4 low bits of a packed face neighbour hold: the
    # neighbour ids are stored unpacked and labels come from the element
    # types; 40 one-millimetre layers with their own (mua, mus, g, n), so
    # photons refract and reflect at every layer boundary
    node, elem = M.meshgen.t6_lattice(30, 30, 40, 1.0)
    c = M.meshgen.centroids(node, elem)
    et = (np.floor(c[:, 2]).astype(np.int32) + 1)
    rng = np.random.default_rng(12)
    med = np.zeros((41, 4), np.float32)
    med[0] = [0, 0, 1, 1]
    for k in range(1, 41):
        med[k] = [rng.uniform(0.001, 0.05), rng.uniform(2, 10), rng.uniform(0.7, 0.95), rng.uniform(1.33, 1.45)]
    tm = M.TetMesh(node, elem, et, med).prepare(nodal=False)
    om = O.OracleMesh(node, elem, et, med, nodal=False)
    det = np.array([[20.0, 15.0, 0.0, 2.0], [10.0, 15.0, 0.0, 2.0]], np.float32)
    cfg = M.SimConfig(nphoton=5000, method="S", srcpos=(15.3, 15.7, 0.0), srcdir=(0.0, 0.0, 1.0), tend=5e-9,
                      tstep=5e-10, issavedet=1, detpos=det, maxdetphoton=5000, seed=77)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=seeds(cfg), records=(0, cfg.nphoton), maxdet=cfg.nphoton, math="portable")
    ok = compare_photons(g.photons, o.photons)
    assert ok.all()
    assert g.detected == o.detected > 0
    gi, oi = np.argsort(g.detphoton), np.argsort(o.detphoton)
    assert np.array_equal(g.detdata[gi], o.detdata[oi])
    assert abs(g.absorbweight - o.absorbweight) <= 1e-6 * o.launchweight
    # and against the reference's arithmetic (the host C library's libm)
    o = O.run(om, cfg, photonseed=seeds(cfg), records=(0, cfg.nphoton), maxdet=cfg.nphoton, math="glibc")
    assert compare_photons(g.photons, o.photons).mean() >= 0.999
    assert abs(g.detected - o.detected) <= 0.001 * cfg.nphoton


def test_invalid_settings_fail_with_a_message(slab):
    # settings the transport cannot honour are refused before any launch,
    # with the reason in mmc_hip_last_error (no silent fallback)
    _, tm, _ = slab
    cube = M.TetMesh(*M.meshgen.cube60(6.0)).prepare(nodal=True)
    cases = [
        (tm, dict(method="G", gridstep=0.0), "voxel edge"),
        (tm, dict(method="G", gridstep=1e-4), "too large"),
        (tm, dict(tstart=0.0, tend=6e-9, tstep=1e-12), "exceeds 2"),     # 375k elements x 6000 gates
        (cube, dict(method="P", omega=1e9), "RF"),
        (tm, dict(omega=-1.0), "omega"),
        (tm, dict(srctype="pattern", srcparam1=(10, 0, 0, 4), srcparam2=(0, 10, 0, 4),
                  srcpattern=np.ones((9, 4, 4), np.float32)), "srcnum"),
        (tm, dict(isreflect=7), "boundary"),
        (tm, dict(srctype="cone", srcparam1=(0.01, 0, 0, 0)), "half angle"),
        (tm, dict(srctype="pencilarray", srcparam1=(10, 0, 0, 1), srcparam2=(0, 10, 0, 3)), "pencil-array"),
        (tm, dict(srctype="pattern3d", srcparam1=(4, 4, 4, 0)), "3-D pattern"),
    ]
    for mesh, kw, msg in cases:
        base = dict(nphoton=100, srcpos=(50.0, 50.0, 0.0) if mesh is tm else (30.0, 30.0, 0.0),
                    srcdir=(0.0, 0.0, 1.0))
        base.update(kw)
        with pytest.raises((RuntimeError, ValueError), match=msg):
            M.run(mesh, M.SimConfig(**base), want_field=False)
    # a batch cannot mix tracers or modulation frequencies
    rm = M.ResidentMesh(tm)
    try:
        with pytest.raises(RuntimeError, match="ray-tracer"):
            rm.run_batch([W.slab_sim(100), M.SimConfig(nphoton=100, method="G", srcpos=(50.0, 50.0, 0.0))])
        with pytest.raises(RuntimeError, match="omega"):
            rm.run_batch([M.SimConfig(nphoton=100, omega=1e9, srcpos=(50.0, 50.0, 0.0)),
                          M.SimConfig(nphoton=100, omega=2e9, srcpos=(50.0, 50.0, 0.0))])
    finally:
        rm.release()


def test_zero_photons_plain_batched_and_served(slab):
    """Empty inputs: a simulation of zero photons returns an empty field and
    zero energies through the plain path; inside a batch it leaves the other
    simulations as they run alone; through the server it completes (with an
    empty field) beside requests that do run."""
    import copy
    from mmc_amd.serve import Request, Server
    _, mesh, _ = slab
    sims = W.slab_batch_sims(3, seed=5, lo=3000, hi=6000)
    empty = copy.copy(sims[1])
    empty.nphoton = 0
    print("case: zero photons, plain path", flush=True)
    r = M.run(mesh, empty)
    assert r.launchweight == 0 and r.absorbweight == 0 and r.raytet == 0 and r.detected == 0
    assert not r.field.any()

    print("case: zero photons inside a batch of three", flush=True)
    t = M.sim.tune_struct(compensated=1)
    rm = M.ResidentMesh(mesh)
    batch = rm.run_batch([sims[0], empty, sims[2]], tune=t)
    alone = [rm.run_batch([s], tune=t)[0] for s in (sims[0], sims[2])]
    rm.release()
    assert batch[1].launchweight == 0 and batch[1].raytet == 0 and not batch[1].field.any()
    for b, a in zip((batch[0], batch[2]), alone):
        assert b.launchweight == a.launchweight and b.raytet == a.raytet
        assert abs(b.absorbweight - a.absorbweight) <= 1e-6 * a.launchweight
        assert np.linalg.norm(b.field - a.field) <= 1e-6 * np.linalg.norm(a.field)

    print("case: zero photons through the server", flush=True)
    reqs = [Request(rid=0, mesh="slab", cfg=sims[0], arrival=0.0, want_field=True),
            Request(rid=1, mesh="slab", cfg=empty, arrival=0.0, want_field=True),
            Request(rid=2, mesh="slab", cfg=sims[2], arrival=0.001, want_field=True)]
    srv = Server({"slab": mesh}, arena_bytes=1 << 27, slice_photons=4000, compensated=True, policy="srpt")
    stats = srv.serve(reqs)
    srv.close()
    assert stats["requests"] == 3
    assert reqs[1].finished >= 0 and reqs[1].launches == 0 and reqs[1].fieldsum == 0 and not reqs[1].field.any()
    for q, a in zip((reqs[0], reqs[2]), alone):
        assert q.launchweight == a.launchweight and q.raytet == a.raytet
        assert np.linalg.norm(q.field - a.field) <= 1e-6 * np.linalg.norm(a.field)
