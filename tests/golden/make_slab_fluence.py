"""Writes slab_fluence_1e7.npz: the time-resolved fluence of the 1 mm,
3.0M-tet four-layer slab (100x100x50 mm, pencil at the top centre, 50 gates
of 0.1 ns) from the oracle built with the host C library's libm -- the
reference CPU tracer's arithmetic -- at 1e7 photons, twice: with the photon
streams of seed 20250101 (the ones the GPU test replays) and of seed
20250102 (an independent run, for the Monte Carlo spread).

The raw field (energy deposits per element and gate) is summed into 5 mm
cubes by element centroid (20x20x10 blocks) to keep the fixture small.

    python tests/golden/make_slab_fluence.py [--photons 1e7] [--threads N]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as O  # noqa: E402

import mmc_amd as M  # noqa: E402
from mmc_amd import workloads as W  # noqa: E402

SEEDS = (20250101, 20250102)
BLOCK = 5.0


def block_index(node, elem):
    """Block of each element by centroid: x-fastest over 20x20x10 blocks."""
    c = M.meshgen.centroids(node, elem)
    b = np.floor(c / BLOCK).astype(np.int64)
    return b[:, 0] + 20 * (b[:, 1] + 20 * b[:, 2])


def blocks(field, bidx):
    """(gates, ne) field -> (gates, 4000) block sums."""
    return np.stack([np.bincount(bidx, weights=g, minlength=4000) for g in field])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--photons", type=float, default=1e7)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(HERE, "slab_fluence_1e7.npz"))
    a = ap.parse_args()
    node, elem, et, med = M.meshgen.slab_layers(1.0)
    om = O.OracleMesh(node, elem, et, med, nodal=False)
    bidx = block_index(node, elem)
    n = int(a.photons)
    out = {}
    for tag, seed in zip("ab", SEEDS):
        cfg = W.slab_sim(n, seed=seed)
        cfg.issavedet, cfg.detpos = 0, None
        t = time.time()
        r = O.run(om, cfg, photonseed=M.photon_seeds(seed, 0, n), nthreads=a.threads, math="glibc")
        print(f"seed {seed}: {n} photons in {time.time() - t:.0f} s, absorbed {r.absorbweight:.6f}", flush=True)
        out[f"field_{tag}"] = blocks(r.field, bidx)
        out[f"absorb_{tag}"] = r.absorbweight
        out[f"seed_{tag}"] = seed
    np.savez_compressed(a.out, photons=n, block=BLOCK, **out)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
