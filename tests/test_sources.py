"""Every source type of launchphoton (src/mmc_raytrace.c:2346-2592).

The oracle's launch is pinned on a clear box (mua = mus = 0, index matched,
no reflection): photons fly straight, so the exit record shows the launch
direction, and, for a source on the bottom face pointing up, the launch
position (x, y) and, through the time of flight, its depth.  Each source is
checked against what its definition fixes: the area of a disk or ring is
covered uniformly (r^2 uniform), a Gaussian beam has r^2 exponential with
mean waist^2, isotropic directions have a uniform cosine, a cone stays
inside its half angle, the arcsine zenith angle is uniform, Fourier weights
are the cosine pattern at the launch position, a pencil array lands on its
grid, a 3-D pattern weights by the voxel the photon starts in, a line emits
across itself, a hyperboloid ray passes the waist plane at its waist
radius, and a focused beam passes through its focus.  The GPU then matches
the oracle photon by photon for every type on a scattering cube.
"""
import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import meshgen

C0 = 299792458e3                      # mm/s (1 / R_C0, src/mmc_const.h:45)
N = 6000


def clear_box():
    node, elem = meshgen.t6_lattice(20, 20, 10, 1.0, origin=(-10.0, -10.0, 0.0))
    etype = np.ones(len(elem), np.int32)
    med = np.array([[0, 0, 1, 1], [0.0, 0.0, 1.0, 1.0]], np.float32)
    c = meshgen.centroids(node, elem)
    return node, elem, etype, med, c


def run_clear(srctype, srcpos=(0.0, 0.0, 0.0), srcdir=(0.0, 0.0, 1.0), p1=(0, 0, 0, 0), p2=(0, 0, 0, 0),
              pattern=None, zrange=(0.0, 1.0), n=N, allow_timeout=False):
    node, elem, etype, med, c = clear_box()
    srcelem = (np.where((c[:, 2] > zrange[0]) & (c[:, 2] < zrange[1]))[0] + 1).astype(np.int32)
    om = O.OracleMesh(node, elem, etype, med, nodal=False, isreflect=0, srcelem=srcelem)
    cfg = M.SimConfig(nphoton=n, seed=5, method="S", srctype=srctype, srcpos=srcpos, srcdir=srcdir,
                      srcparam1=p1, srcparam2=p2, srcpattern=pattern, tstart=0.0, tend=1e-8, tstep=1e-8,
                      outputtype="f", isreflect=0, minenergy=0.0, nout=1.0)
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, n), records=(0, n), math="portable")
    ph = o.photons
    if allow_timeout:                         # slow photons may reach tend first
        ph = ph[ph["reason"] == 1]
        assert len(ph) > 0.99 * n
    assert (ph["reason"] == 1).all()          # every photon leaves through the box's surface
    return ph


def within(mean, want, sd, n, k=5.0):
    return abs(mean - want) <= k * sd / np.sqrt(n)


def test_disk_ring_gaussian_cover_their_areas():
    ph = run_clear("disk", p1=(3.0, 0, 0, 0))
    r2 = (ph["p"][:, :2].astype(np.float64) ** 2).sum(1)
    assert np.allclose(ph["p"][:, 2], 10.0, atol=1e-4) and r2.max() <= 9.0 + 1e-4
    assert within(r2.mean(), 4.5, 9.0 / np.sqrt(12), N)                     # r^2 ~ U(0, 9)
    ph = run_clear("ring", p1=(3.0, 2.0, 0, 0))
    r2 = (ph["p"][:, :2].astype(np.float64) ** 2).sum(1)
    assert r2.min() >= 4.0 - 1e-3 and r2.max() <= 9.0 + 1e-3
    assert within(r2.mean(), 6.5, 5.0 / np.sqrt(12), N)
    # a ring sector: azimuth within [z, w]
    ph = run_clear("ring", p1=(3.0, 2.0, 0.5, 1.5))
    phi = np.arctan2(ph["p"][:, 1], ph["p"][:, 0])
    assert phi.min() >= 0.5 - 1e-4 and phi.max() <= 1.5 + 1e-4
    ph = run_clear("gaussian", p1=(2.0, 0, 0, 0))
    r2 = (ph["p"][:, :2].astype(np.float64) ** 2).sum(1)
    assert within(r2.mean(), 4.0, 4.0, N)                                    # r^2 ~ Exp(waist^2)


def test_direction_sources():
    for st, p1 in (("isotropic", (0, 0, 0, 0)), ("cone", (0.5, 0, 0, 0)), ("cone", (0.5, 1, 0, 0)),
                   ("arcsine", (0, 0, 0, 0))):
        ph = run_clear(st, srcpos=(0.0, 0.0, 5.0), p1=p1)
        v = ph["v"].astype(np.float64)
        assert np.allclose((v ** 2).sum(1), 1.0, atol=1e-5)
        th = np.arccos(np.clip(v[:, 2], -1, 1))
        if st == "isotropic":
            assert within(v[:, 2].mean(), 0.0, np.sqrt(1 / 3), N)
            assert within((v[:, 2] ** 2).mean(), 1 / 3, np.sqrt(4 / 45), N)
        elif st == "cone" and p1[1] == 0:
            ca = np.cos(0.5)
            assert v[:, 2].min() >= ca - 1e-5
            assert within(v[:, 2].mean(), (1 + ca) / 2, (1 - ca) / np.sqrt(12), N)   # cos uniform
        elif st == "cone":
            assert th.max() <= 0.5 + 1e-4
            assert within(th.mean(), 0.25, 0.5 / np.sqrt(12), N)                   # angle uniform
        else:
            assert within(th.mean(), np.pi / 2, np.pi / np.sqrt(12), N)
        assert within(np.arctan2(v[:, 1], v[:, 0]).mean(), 0.0, np.pi / np.sqrt(3), N)


def test_zgaussian_angles():
    s = 0.2
    ph = run_clear("zgaussian", srcpos=(0.0, 0.0, 5.0), p1=(s, 0, 0, 0))
    th = np.arccos(np.clip(ph["v"][:, 2].astype(np.float64), -1, 1))
    # zenith sqrt(-2 ln u1) (1 - 2 u2) sigma: E[theta^2] = 2 sigma^2 / 3
    assert within((th ** 2).mean(), 2 * s * s / 3, 2 * s * s, N)


def test_weighted_plane_sources():
    # Fourier: w = (cos(2 pi (kx rx + ky ry + phase)) (1 - m) + 1) / 2
    ph = run_clear("fourier", srcpos=(-5.0, -5.0, 0.0), p1=(10, 0, 0, 2.25), p2=(0, 10, 0, 1.5))
    rx, ry = (ph["p"][:, 0] + 5.0) / 10, (ph["p"][:, 1] + 5.0) / 10
    want = (np.cos(2 * np.pi * (2 * rx + 1 * ry + 0.25)) * (1 - 0.5) + 1) / 2
    assert np.allclose(ph["w0"], want, rtol=0, atol=1e-3)   # exit vs launch: face nudges
    # fourierx: second axis = w * (dir x param1)/|param1|
    ph = run_clear("fourierx", srcpos=(-5.0, -5.0, 0.0), p1=(10, 0, 0, 1.0), p2=(2, 1, 0.25, 0.5))
    rx, ry = (ph["p"][:, 0] + 5.0) / 10, (ph["p"][:, 1] + 5.0) / 1.0
    assert ry.min() >= -1e-4 and ry.max() <= 1 + 1e-4
    want = (np.cos(2 * np.pi * (2 * rx + 1 * ry + 0.25)) * (1 - 0.5) + 1) / 2
    assert np.allclose(ph["w0"], want, rtol=0, atol=1e-3)   # exit vs launch: face nudges
    ph = run_clear("fourierx2d", srcpos=(-5.0, -5.0, 0.0), p1=(10, 0, 0, 10.0), p2=(2, 3, 0.1, 0.2))
    rx, ry = (ph["p"][:, 0] + 5.0) / 10, (ph["p"][:, 1] + 5.0) / 10
    want = (np.sin(2 * np.pi * (2 * rx + 0.1)) * np.sin(2 * np.pi * (3 * ry + 0.2)) + 1) / 2
    assert np.allclose(ph["w0"], want, rtol=0, atol=1e-3)   # exit vs launch: face nudges


def test_pencil_array_and_pattern3d():
    ph = run_clear("pencilarray", srcpos=(-4.3, -3.3, 0.0), p1=(8, 0, 0, 5), p2=(0, 6, 0, 4))
    gx, gy = -4.3 + 2.0 * np.arange(5), -3.3 + 2.0 * np.arange(4)
    ix = np.abs(ph["p"][:, 0, None] - gx).argmin(1)
    iy = np.abs(ph["p"][:, 1, None] - gy).argmin(1)
    assert np.abs(ph["p"][:, 0] - gx[ix]).max() < 1e-3 and np.abs(ph["p"][:, 1] - gy[iy]).max() < 1e-3
    counts = np.bincount(iy * 5 + ix, minlength=20)
    assert counts.min() > N / 20 * 0.7                   # every beam of the 5 x 4 array
    # 3-D pattern in a 4x4x4 mm box: weight of the starting voxel (z from the time of flight)
    pat = (1.0 + np.arange(64, dtype=np.float32)).reshape(4, 4, 4) / 64
    ph = run_clear("pattern3d", srcpos=(-2.0, -2.0, 1.0), p1=(4, 4, 4, 0), pattern=pat, zrange=(0.0, 6.0))
    z0 = 10.0 - ph["t"].astype(np.float64) * C0
    ix = np.floor((ph["p"][:, 0] + 2.0) * 0.9998).astype(int)
    iy = np.floor((ph["p"][:, 1] + 2.0) * 0.9998).astype(int)
    iz = np.floor((z0 - 1.0) * 0.9998).astype(int)
    frac = np.stack([(ph["p"][:, 0] + 2.0) % 1, (ph["p"][:, 1] + 2.0) % 1, (z0 - 1.0) % 1])
    clear = ((frac > 0.01) & (frac < 0.99)).all(0)
    assert clear.sum() > N // 2
    assert np.array_equal(ph["w0"][clear], pat[iz, iy, ix][clear])


def test_shared_patterns_weight_each_pattern():
    # photon sharing (srcnum > 1, src/mmc_raytrace.c:410-423, 1586-1612,
    # 1812-1835) in an absorbing box without scattering: each photon flies
    # straight up 10 mm, so its launch pixel is read back from its exit
    # (x, y), and pattern k's launched energy is the sum of its values at
    # those pixels and its absorbed energy that sum times 1 - exp(-10 mua).
    # x and y do not change on the way up, so the pixel is exact away from
    # its edges.  Slot k of the field (energy / mua per element for fluence
    # output) equals the run with pattern k alone.
    node, elem, etype, med, c = clear_box()
    mua = 0.05
    med[1] = (mua, 0.0, 1.0, 1.0)
    srcelem = (np.where(c[:, 2] < 1.0)[0] + 1).astype(np.int32)
    om = O.OracleMesh(node, elem, etype, med, nodal=False, isreflect=0, srcelem=srcelem)
    pats = np.random.default_rng(9).uniform(0.2, 1.0, (3, 4, 5)).astype(np.float32)
    kw = dict(nphoton=N, seed=5, method="S", srctype="pattern", srcpos=(-8.0, -6.0, 0.0), srcdir=(0.0, 0.0, 1.0),
              srcparam1=(15.0, 0, 0, 5), srcparam2=(0, 12.0, 0, 4), tstart=0.0, tend=1e-8, tstep=1e-8,
              outputtype="f", isreflect=0, minenergy=0.0, nout=1.0)
    seeds = M.photon_seeds(5, 0, N)
    sh = O.run(om, M.SimConfig(srcpattern=pats, **kw), photonseed=seeds, records=(0, N), math="portable")
    ph = sh.photons
    assert (ph["reason"] == 1).all() and np.allclose(ph["p"][:, 2], 10.0, atol=1e-4)
    fx, fy = (ph["p"][:, 0] + 8.0) / 3.0, (ph["p"][:, 1] + 6.0) / 3.0
    clear = (np.abs(fx - np.round(fx)) > 1e-5) & (np.abs(fy - np.round(fy)) > 1e-5)
    assert clear.all()                                   # no photon on a pixel edge at this seed
    ix, iy = np.floor(fx).astype(int), np.floor(fy).astype(int)
    assert sh.field.shape == (1, om.ne, 3) and sh.srcweight.shape == (2, 3)
    for k in range(3):
        w = pats[k, iy, ix].astype(np.float64)
        assert abs(sh.srcweight[0, k] / w.sum() - 1) < 1e-6
        assert abs(sh.srcweight[1, k] / (w.sum() * -np.expm1(-10 * mua)) - 1) < 1e-5
        assert abs(sh.field[..., k].sum() * mua / sh.srcweight[1, k] - 1) < 1e-5   # fluence: energy / mua
        one = O.run(om, M.SimConfig(srcpattern=pats[k], **kw), photonseed=seeds, math="portable")
        assert abs(one.launchweight / sh.srcweight[0, k] - 1) < 1e-6
        assert np.abs(sh.field[..., k] - one.field).max() <= 1e-6 * np.abs(one.field).max()


def test_line_slit_hyperboloid_focus():
    ph = run_clear("line", srcpos=(-3.0, 0.0, 5.0), srcdir=(1.0, 0.0, 0.0), p1=(6, 0, 0, 0), zrange=(4.0, 6.0),
                   allow_timeout=True)
    v = ph["v"].astype(np.float64)
    # srcdir x (t, s, p) with p = +-sqrt(1 - vx^2 - vy^2) = 0 here: (0, 0, s), s ~ U(-1, 1),
    # left unnormalized as the reference does (src/mmc_raytrace.c:2552-2562)
    assert np.abs(v[:, :2]).max() < 1e-6
    assert within(v[:, 2].mean(), 0.0, 1 / np.sqrt(3), N) and within(np.abs(v[:, 2]).mean(), 0.5, 1 / np.sqrt(12), N)
    assert np.abs(v[:, 2]).min() < 0.01        # far from unit length for some photons
    # the step length is measured in units of the direction vector, so a photon
    # starting 5 mm from the top and bottom faces flies for 5 / |s| mm of "path"
    assert np.allclose(ph["t"].astype(np.float64) * C0, 5.0 / np.abs(v[:, 2]), rtol=5e-4)
    ph = run_clear("slit", srcpos=(-3.0, 0.3, 0.0), p1=(6, 0, 0, 0))
    assert np.allclose(ph["p"][:, 1], 0.3, atol=1e-4) and ph["p"][:, 0].min() >= -3 - 1e-4
    assert np.allclose(ph["v"][:, 2], 1.0, atol=1e-6)
    # hyperboloid: waist w at z = 4 (srcparam1 = w, waist distance, Rayleigh length)
    w, zw, zr = 1.5, 4.0, 3.0
    ph = run_clear("hyperboloid", p1=(w, zw, zr, 0))
    p, v = ph["p"].astype(np.float64), ph["v"].astype(np.float64)
    at = lambda z: p[:, :2] + ((z - p[:, 2]) / v[:, 2])[:, None] * v[:, :2]
    r_w = np.sqrt((at(zw) ** 2).sum(1))
    assert np.all(r_w <= np.sqrt((at(0.0) ** 2).sum(1)) + 1e-4)
    assert np.all(r_w <= np.sqrt((at(8.0) ** 2).sum(1)) + 1e-4)
    # 0.5 * scattering length: r^2 ~ Exp(w^2 / 2)
    assert within((r_w ** 2).mean(), w * w / 2, w * w / 2, N)
    # a planar beam focused 6 mm above its centre
    ph = run_clear("planar", srcpos=(-2.0, -2.0, 0.0), srcdir=(0, 0, 1, 6.0), p1=(4, 0, 0, 0), p2=(0, 4, 0, 0))
    p, v = ph["p"].astype(np.float64), ph["v"].astype(np.float64)
    f = np.array([0.0, 0.0, 6.0])
    miss = np.linalg.norm(np.cross(f - p, v), axis=1)
    assert miss.max() < 1e-3


CASES = {
    "isotropic": dict(p1=(0, 0, 0, 0)), "cone": dict(p1=(0.6, 0, 0, 0)), "arcsine": dict(p1=(0, 0, 0, 0)),
    "gaussian": dict(p1=(3.0, 0, 0, 0)), "disk": dict(p1=(4.0, 0, 0, 0)), "ring": dict(p1=(4.0, 2.0, 0, 0)),
    "fourier": dict(p1=(12, 0, 0, 2.25), p2=(0, 12, 0, 1.5), corner=True),
    "fourierx": dict(p1=(12, 0, 0, 12), p2=(2, 1, 0.25, 0.5), corner=True),
    "fourierx2d": dict(p1=(12, 0, 0, 12), p2=(2, 3, 0.1, 0.2), corner=True),
    "zgaussian": dict(p1=(0.3, 0, 0, 0)), "line": dict(p1=(10, 0, 0, 0), dir=(1.0, 0.0, 0.0)),
    "slit": dict(p1=(10, 0, 0, 0), shift=True), "pencilarray": dict(p1=(12, 0, 0, 4), p2=(0, 12, 0, 3), corner=True),
    "pattern3d": dict(p1=(8, 8, 3, 0), corner=True), "hyperboloid": dict(p1=(2.0, 5.0, 4.0, 0)),
    "planar-focus": dict(p1=(12, 0, 0, 0), p2=(0, 12, 0, 0), corner=True, focus=-20.0),
}


def source_case(name):
    """The scattering cube (4 mm cells, source elements below z = 8 mm) and
    the simulation of source case `name`: (device mesh, oracle mesh, cfg)."""
    kw = CASES[name]
    st = "planar" if name == "planar-focus" else name
    node, elem, et, med = meshgen.cube60(4.0)
    c = meshgen.centroids(node, elem)
    srcelem = (np.where(c[:, 2] < 8.0)[0] + 1).astype(np.int32)
    tm = M.TetMesh(node, elem, et, med, srcelem=srcelem).prepare(nodal=False)
    om = O.OracleMesh(node, elem, et, med, nodal=False, srcelem=srcelem)
    pos = (24.3, 24.7, 0.0) if kw.get("corner") else (30.0, 30.0, 0.0)
    if kw.get("shift"):
        pos = (25.0, 30.0, 0.0)
    if st == "line":
        pos = (25.0, 30.0, 2.0)
    if st in ("isotropic", "cone", "arcsine", "zgaussian"):
        pos = (30.3, 30.7, 5.1)          # off the lattice planes
    d = kw.get("dir", (0.0, 0.0, 1.0))
    cfg = M.SimConfig(nphoton=6000, method="S", srctype=st, srcpos=pos, srcdir=tuple(d) + (kw.get("focus", 0.0),),
                      srcparam1=kw["p1"], srcparam2=kw.get("p2", (0, 0, 0, 0)), tend=5e-9, tstep=5e-10, seed=23,
                      srcpattern=(np.random.default_rng(4).random((3, 8, 8)).astype(np.float32)
                                  if st == "pattern3d" else None))
    return tm, om, cfg


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_sources_per_photon(name):
    tm, om, cfg = source_case(name)
    print(f"source {name}", flush=True)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              math="portable")
    same = (g.photons["eid"] == o.photons["eid"]) & (g.photons["reason"] == o.photons["reason"])
    dp = np.abs(g.photons["p"].astype(np.float64) - o.photons["p"]).max(axis=1)
    ok = same & (dp <= 1e-4)
    print(f"{name}: same {ok.mean():.5f}, w0 max diff {np.abs(g.photons['w0'] - o.photons['w0']).max():.3g}",
          flush=True)
    assert ok.mean() >= 0.999
    assert np.array_equal(g.photons["w0"], o.photons["w0"])
    # the oracle stops at a launch that finds no element; on the GPU every photon was traced
    assert (g.photons["nraytet"] > 0).all()
    assert g.launchweight == pytest.approx(o.launchweight, rel=1e-6)
