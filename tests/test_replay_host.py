"""Seed replay on the host and in the oracle: the .mch seed block, the
vectorized per-photon seeds, the restatement of mesh_loadseedfile
(src/mmc_mesh.c:815-898), and the replay outputs of the oracle against what
they must sum to (src/mmc_raytrace.c:1523-1557)."""
import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import replay as R
from mmc_amd import workloads as W


def test_seeds_of_equals_the_library_rule():
    a = R.seeds_of(0x623F9A9E, np.arange(5000, 9000))
    assert np.array_equal(a, M.photon_seeds(0x623F9A9E, 5000, 4000))
    ids = np.array([3, 1_000_000_007, 42])
    assert np.array_equal(R.seeds_of(7, ids)[1], M.photon_seeds(7, 1_000_000_007, 1)[0])


class _Res:
    def __init__(self, rec, ids, detected):
        self.detdata, self.detphoton, self.detected = rec, ids, detected


def test_mch_round_trip_with_seeds(tmp_path):
    rec = np.arange(4 * 6, dtype=np.float32).reshape(4, 6)
    cfg = M.SimConfig(nphoton=100, seed=9, detpos=np.zeros((2, 4), np.float32), issavedet=1)
    res = _Res(rec, np.array([5, 17, 33, 80]), 4)
    seeds = R.detected_seeds(cfg, res)
    R.save_mch(str(tmp_path / "a.mch"), 2, cfg, res, unitinmm=0.5, seeds=seeds)
    f = R.load_mch(str(tmp_path / "a.mch"))
    assert (f.maxmedia, f.detnum, f.colcount, f.totalphoton, f.detected) == (2, 2, 6, 100, 4)
    assert f.unitinmm == 0.5
    assert np.array_equal(f.records, rec) and np.array_equal(f.seeds, seeds)
    assert (tmp_path / "a.mch").stat().st_size == 64 + 4 * 4 * 6 + 16 * 4


def test_replay_config_weights_times_and_detector_filter():
    # one medium: columns [detid, nscat, ppath, w0]; mesh_loadseedfile uses
    # columns 2..maxmedia+1 with media 1..maxmedia
    med = np.array([[0, 0, 1, 1], [0.02, 1.0, 0.9, 1.4]], np.float32)
    rec = np.array([[1, 12, 30.0, 1.0], [2, 40, 55.5, 0.5], [1, 3, 7.25, 1.0]], np.float32)
    f = R.MchFile(1, 2, 4, 1000, 3, 1.0, 1.0, rec, np.arange(6, dtype=np.uint64).reshape(3, 2))
    cfg = M.SimConfig(nphoton=1000, minenergy=1e-6)
    r = R.replay_config(cfg, f, med, "l")
    assert r.nphoton == 3 and r.minenergy == 0.0
    assert np.allclose(r.replayweight, rec[:, 3] * np.exp(-0.02 * rec[:, 2]), rtol=1e-6)
    assert np.allclose(r.replaytime, 1.4 * rec[:, 2] * 3.335640951981520e-12, rtol=1e-6)
    r1 = R.replay_config(cfg, f, med, "x", replaydet=1)
    assert r1.nphoton == 2 and np.array_equal(r1.photonseed, f.seeds[[0, 2]])
    with pytest.raises(ValueError):
        R.replay_config(cfg, R.MchFile(1, 2, 4, 1000, 3, 1.0, 1.0, rec, None), med)


@pytest.fixture(scope="module")
def detected():
    arrays = M.meshgen.slab_layers(4.0)
    om = O.OracleMesh(*arrays, nodal=False)
    cfg = W.slab_sim(30000, seed=17)
    seeds = M.photon_seeds(cfg.seed, 0, cfg.nphoton)
    o = O.run(om, cfg, photonseed=seeds, records=(0, cfg.nphoton), maxdet=cfg.nphoton, math="portable")
    assert o.detected > 100
    return arrays, om, cfg, o, seeds


def test_oracle_replay_reproduces_each_detected_photon(detected):
    arrays, om, cfg, o, seeds = detected
    ids = np.sort(o.detphoton)
    order = np.argsort(o.detphoton)
    rcfg = M.SimConfig(**{**cfg.__dict__, "nphoton": len(ids)})
    r = O.run(om, rcfg, photonseed=seeds[ids], records=(0, len(ids)), maxdet=len(ids), math="portable")
    assert r.detected == len(ids)
    assert np.array_equal(r.photons["p"], o.photons["p"][ids])
    ri = np.argsort(r.detphoton)
    assert np.array_equal(r.detdata[ri], o.detdata[order])


def test_oracle_weighted_path_and_scattering_outputs(detected):
    arrays, om, cfg, o, seeds = detected
    ids = np.sort(o.detphoton)
    prop = len(arrays[3]) - 1
    rec = o.detdata[np.argsort(o.detphoton)]
    ppath = rec[:, 1 + prop:1 + 2 * prop].astype(np.float64).sum(axis=1)
    rw = np.random.default_rng(2).uniform(0.5, 2.0, len(ids)).astype(np.float32)
    base = {**cfg.__dict__, "nphoton": len(ids), "issavedet": 0, "minenergy": 0.0,
            "replayweight": rw, "replaytime": np.full(len(ids), 1e-10, np.float32)}
    wl = O.run(om, M.SimConfig(**{**base, "outputtype": "l"}), photonseed=seeds[ids], math="portable")
    # every step deposits L * rw: the total is the detected photons' paths
    assert abs(wl.field.sum() / (rw * ppath).sum() - 1) < 1e-5
    # all of it in the gate of the detected time
    assert np.count_nonzero(wl.field.sum(axis=1)) == 1 and wl.field[1].sum() > 0
    assert wl.launchweight == pytest.approx(float(rw.astype(np.float64).sum()), rel=1e-6)
    wp = O.run(om, M.SimConfig(**{**base, "outputtype": "p"}), photonseed=seeds[ids], math="portable")
    nscat = (o.photons["nscat"][ids] * rw).sum()
    steps = (o.photons["nraytet"][ids] * rw).sum()
    # each scattering closes a segment whose fractions L*mus/slen0 add to 1;
    # steps before the first scattering count 1 each, no step more than 1
    assert nscat * (1 - 1e-4) <= wp.field.sum() <= steps


def test_oracle_pattern_replay_weights_take_the_launch_pixel():
    """Seed replay of the weighted path output with a pattern source: each
    detected photon's replay weight is scaled by its launch pixel's pattern
    value (src/mmc_raytrace.c:2394-2399), so the L output sums to
    sum(rw * pattern(pixel) * path) and the launched energy to
    sum(rw * pattern(pixel)); the pattern value is the record's launch weight."""
    arrays = M.meshgen.slab_layers(4.0)
    c = M.meshgen.centroids(arrays[0], arrays[1])
    srcelem = (np.where(c[:, 2] < 8.0)[0] + 1).astype(np.int32)      # launch candidates
    om = O.OracleMesh(*arrays, nodal=False, srcelem=srcelem)
    pat = np.random.default_rng(8).uniform(0.2, 1.0, (8, 8)).astype(np.float32)
    cfg = W.slab_sim(30000, seed=41)
    cfg.srctype, cfg.srcpos = "pattern", (30.3, 30.7, 0.0)
    cfg.srcparam1, cfg.srcparam2, cfg.srcpattern = (40.0, 0.0, 0.0, 8), (0.0, 40.0, 0.0, 8), pat
    seeds = M.photon_seeds(cfg.seed, 0, cfg.nphoton)
    o = O.run(om, cfg, photonseed=seeds, records=(0, cfg.nphoton), maxdet=cfg.nphoton, math="portable")
    assert o.detected > 100
    ids = np.sort(o.detphoton)
    prop = len(arrays[3]) - 1
    rec = o.detdata[np.argsort(o.detphoton)]
    ppath = rec[:, 1 + prop:1 + 2 * prop].astype(np.float64).sum(axis=1)
    w0 = rec[:, -1].astype(np.float64)                      # the launch pixel's pattern value
    assert len(np.unique(w0)) > 5 and np.all(np.isin(w0.astype(np.float32), pat))
    rw = np.random.default_rng(3).uniform(0.5, 2.0, len(ids)).astype(np.float32)
    base = {**cfg.__dict__, "nphoton": len(ids), "issavedet": 0, "minenergy": 0.0, "outputtype": "l",
            "replayweight": rw, "replaytime": np.full(len(ids), 1e-10, np.float32)}
    wl = O.run(om, M.SimConfig(**base), photonseed=seeds[ids], math="portable")
    assert abs(wl.field.sum() / (rw * w0 * ppath).sum() - 1) < 1e-5
    assert wl.launchweight == pytest.approx(float((rw.astype(np.float64) * w0).sum()), rel=1e-6)


def _chords(node, elem, p0, d, tmax):
    """FP64 length of the ray p0 + t d, 0 <= t <= tmax, inside each tetrahedron
    (intersection of the four half-spaces, brute force over all elements)."""
    P = node[elem - 1].astype(np.float64)                          # (ne, 4, 3)
    lo, hi = np.zeros(len(elem)), np.full(len(elem), tmax)
    for f in range(4):
        a, b, c = (P[:, (f + k) % 4] for k in (1, 2, 3))
        n = np.cross(b - a, c - a)
        n *= np.sign(((P[:, f] - a) * n).sum(1))[:, None]           # inward: towards the fourth node
        num, den = ((a - p0) * n).sum(1), (d * n).sum(1)            # inside: den * t >= num
        with np.errstate(divide="ignore", invalid="ignore"):
            t = num / den
        lo = np.where(den > 0, np.maximum(lo, t), lo)
        hi = np.where(den < 0, np.minimum(hi, t), hi)
        lo = np.where((den == 0) & (num > 0), np.inf, lo)
    return np.maximum(hi - lo, 0.0)


def test_oracle_jacobian_replay_closed_form():
    """J replay of a non-scattering pencil beam (src/mmc_raytrace.c:1523-1527,
    1568-1570, mmc_fluence_deposit at :61-63): a step of length L in an
    element of absorption mua deposits rw e^(-dmua L) / mua, and, where
    mua = 0, (rw e^(-dmua L) + w) L with w the photon's weight there, dmua =
    1e-4 /mm (src/mmc_mesh.h:73).  The chord of the beam in every element
    comes from an FP64 ray-tetrahedron clip over all elements."""
    node, elem = M.meshgen.t6_lattice(8, 8, 10, 1.0, origin=(-4.0, -4.0, 0.0))
    c = M.meshgen.centroids(node, elem)
    etype = np.where(c[:, 2] < 5.0, 1, 2).astype(np.int32)
    mua1 = 0.02
    med = np.array([[0, 0, 1, 1], [mua1, 0.0, 1.0, 1.0], [0.0, 0.0, 1.0, 1.0]], np.float32)
    srcelem = (np.where(c[:, 2] < 1.0)[0] + 1).astype(np.int32)
    om = O.OracleMesh(node, elem, etype, med, nodal=False, isreflect=0, srcelem=srcelem)
    d = np.array([0.15, 0.1, 1.0])
    d /= np.linalg.norm(d)
    p0 = np.array([0.31, -0.17, 0.0])
    L = _chords(node, elem, p0, d, 10.0 / d[2])
    hit = L > 0
    assert L[hit].min() > 1e-3 and abs(L.sum() - 10.0 / d[2]) < 1e-9      # no edge grazed
    n = 16
    rw = np.random.default_rng(4).uniform(0.5, 2.0, n).astype(np.float32)
    cfg = M.SimConfig(nphoton=n, seed=5, method="S", srctype="pencil", srcpos=tuple(p0), srcdir=tuple(d),
                      tstart=0.0, tend=1e-8, tstep=1e-8, outputtype="j", isreflect=0, minenergy=0.0, nout=1.0,
                      replayweight=rw, replaytime=np.full(n, 1e-10, np.float32))
    o = O.run(om, cfg, photonseed=M.photon_seeds(5, 0, n), math="portable")
    f = o.field[0]
    assert np.count_nonzero(f) == hit.sum() and np.all(f[hit] > 0)
    mua = med[etype, 0].astype(np.float64)
    # the weight where the beam enters each element of the clear layer: all
    # of the absorbing layer's 5 / d_z mm behind it
    w = np.where(mua > 0, 0.0, np.exp(-mua1 * 5.0 / d[2]))
    e = np.exp(-1e-4 * L)
    want = np.where(mua > 0, rw.sum(dtype=np.float64) * e / np.where(mua > 0, mua, 1.0),
                    (rw.sum(dtype=np.float64) * e + n * w) * L)
    assert (mua[hit] > 0).sum() > 10 and (mua[hit] == 0).sum() > 10
    assert np.allclose(f[hit], want[hit], rtol=2e-5, atol=0)
