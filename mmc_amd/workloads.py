"""The benchmark workloads as (mesh, settings) pairs.

cube60      homogeneous 60 mm cube, 2 mm cells x6 tets (162,000 tets); Plucker
            nodal fluence, 1 gate 0-5 ns, pencil at (30,30,0) +z
slab        100x100x50 mm box, 1 mm cells x6 tets (3.0M tets), four layers and a
            10 mm sphere; branch-less Badouel element fluence, 50 gates of
            0.1 ns, 16 disk detectors (r=1.5 mm) at 5-40 mm offsets
slab_batch  256 simulations on the slab mesh: pencils on a 16x16 grid of 4 mm
            pitch, per-tissue media drawn from a seed, photon counts
            log-uniform in [1e5, 1e7], 10 gates of 0.5 ns
shells      jittered 200x200x100 mm box, 1 mm cells x6 tets (24.0M tets; the
            stated ~22M tets implies ~1 mm cells, not 0.6 mm which would give
            111M), five tissues in nested ellipsoidal shells; a 60x60 mm planar
            source with a 64x64 Bernoulli(0.5) pattern, 100 gates of 0.05 ns,
            Plucker nodal fluence, 32 detectors; the 8 patterns either share
            every photon (srcnum = 8) or run as 8 pattern requests
"""
from __future__ import annotations

import numpy as np

from . import meshgen
from .mesh import TetMesh
from .sim import SimConfig

CUBE60_SEED = 1648335518


def cube60_mesh(h: float = 2.0) -> TetMesh:
    node, elem, et, med = meshgen.cube60(h)
    return TetMesh(node, elem, et, med).prepare(nodal=True)


def cube60_sim(nphoton: int = 1_000_000, seed: int = CUBE60_SEED) -> SimConfig:
    return SimConfig(nphoton=nphoton, seed=seed, method="P", srcpos=(30.0, 30.0, 0.0), srcdir=(0.0, 0.0, 1.0),
                     tstart=0.0, tend=5e-9, tstep=5e-9, outputtype="x")


def slab_detectors() -> np.ndarray:
    offs = np.linspace(5.0, 40.0, 16)
    return np.stack([50.0 + offs, np.full(16, 50.0), np.zeros(16), np.full(16, 1.5)], 1).astype(np.float32)


def slab_mesh(h: float = 1.0) -> TetMesh:
    node, elem, et, med = meshgen.slab_layers(h)
    return TetMesh(node, elem, et, med).prepare(nodal=False)


def slab_sim(nphoton: int = 100_000_000, seed: int = 20250101) -> SimConfig:
    return SimConfig(nphoton=nphoton, seed=seed, method="S", srcpos=(50.0, 50.0, 0.0), srcdir=(0.0, 0.0, 1.0),
                     tstart=0.0, tend=5e-9, tstep=1e-10, outputtype="x", issavedet=1,
                     detpos=slab_detectors(), maxdetphoton=1_000_000)


def slab_batch_sims(nsim: int = 256, seed: int = 7, prop: int = 5, lo=1e5, hi=1e7):
    """The batched workload: per-simulation source, media and photon count."""
    rng = np.random.default_rng(seed)
    sims = []
    for k in range(nsim):
        i, j = k % 16, k // 16
        x = 50.0 + (i - 7.5) * 4.0
        y = 50.0 + (j - 7.5) * 4.0
        med = np.zeros((prop + 1, 4), np.float32)
        med[0] = [0, 0, 1, 1]
        for t in range(1, prop + 1):
            g = rng.uniform(0.8, 0.95)
            musp = rng.uniform(0.5, 2.0)
            med[t] = [rng.uniform(0.001, 0.05), musp / (1.0 - g), g, rng.uniform(1.33, 1.45)]
        n = int(round(10 ** rng.uniform(np.log10(lo), np.log10(hi))))
        sims.append(SimConfig(nphoton=n, seed=1000 + k, method="S", srcpos=(x, y, 0.0), srcdir=(0.0, 0.0, 1.0),
                              tstart=0.0, tend=5e-9, tstep=5e-10, outputtype="x", med=med))
    return sims


def shells_arrays(h: float = 1.0, window=(70.0, 130.0)):
    """(node, elem, etype, med, srcelem) of the large jittered mesh; srcelem
    are the launch candidates of the planar source: the elements of the top
    cell layer under the source window."""
    n = int(round(200 / h))
    node, elem, et, med = meshgen.shells_box(n, n, n // 2, h)
    c = meshgen.centroids(node, elem)
    lo, hi = window[0] - 2 * h, window[1] + 2 * h
    top = (c[:, 2] < h) & (c[:, 0] > lo) & (c[:, 0] < hi) & (c[:, 1] > lo) & (c[:, 1] < hi)
    srcelem = (np.nonzero(top)[0] + 1).astype(np.int32)
    return node, elem, et, med, srcelem


def shells_mesh(h: float = 1.0, window=(70.0, 130.0)):
    node, elem, et, med, srcelem = shells_arrays(h, window)
    return TetMesh(node, elem, et, med, srcelem=srcelem).prepare(nodal=True)


def shells_detectors() -> np.ndarray:
    ang = np.arange(32) * (2 * np.pi / 32)
    return np.stack([100 + 45 * np.cos(ang), 100 + 45 * np.sin(ang), np.zeros(32), np.full(32, 1.5)],
                    1).astype(np.float32)


def shells_patterns(npat: int = 8, seed: int = 5) -> np.ndarray:
    """The 8 Bernoulli(0.5) 64x64 patterns over the 60x60 mm window."""
    rng = np.random.default_rng(seed)
    return (rng.random((npat, 64, 64)) < 0.5).astype(np.float32)


def shells_shared_sim(photons: int = 1_000_000_000, npat: int = 8, seed: int = 900) -> SimConfig:
    """The large run with its 8 patterns batched onto every photon
    (srcnum = 8, photon sharing): one trajectory serves all patterns and the
    nodal field holds 8 slots per node and gate."""
    return SimConfig(nphoton=photons, seed=seed, method="P", srctype="pattern",
                     srcpos=(70.0, 70.0, 0.0), srcdir=(0.0, 0.0, 1.0),
                     srcparam1=(60.0, 0.0, 0.0, 64.0), srcparam2=(0.0, 60.0, 0.0, 64.0),
                     srcpattern=shells_patterns(npat), tstart=0.0, tend=5e-9, tstep=5e-11, outputtype="x",
                     issavedet=1, detpos=shells_detectors(), maxdetphoton=1_000_000)


def shells_sims(nreq: int = 8, photons: int = 1_000_000_000, seed: int = 5):
    """8 pattern requests sharing the 1e9 photons, each with its own 64x64
    Bernoulli(0.5) pattern over the 60x60 mm window."""
    rng = np.random.default_rng(seed)
    sims = []
    for k in range(nreq):
        pat = (rng.random((64, 64)) < 0.5).astype(np.float32)
        sims.append(SimConfig(nphoton=photons // nreq, seed=900 + k, method="P", srctype="pattern",
                              srcpos=(70.0, 70.0, 0.0), srcdir=(0.0, 0.0, 1.0),
                              srcparam1=(60.0, 0.0, 0.0, 64.0), srcparam2=(0.0, 60.0, 0.0, 64.0),
                              srcpattern=pat, tstart=0.0, tend=5e-9, tstep=5e-11, outputtype="x",
                              issavedet=1, detpos=shells_detectors(), maxdetphoton=1_000_000))
    return sims
