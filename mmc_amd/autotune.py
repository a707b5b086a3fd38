"""Launch-shape autotuner for the persistent transport kernel.

For a batch of simulations on a resident mesh it times short calibration
launches (every simulation scaled to a small photon sample, same sources and
media) over a grid of workgroup size, resident workgroups per CU and photon
ids claimed per lane refill, and keeps the shape with the most ray-tet steps
per millisecond.  Results are cached per (mesh size, ray-tracer, batch size).
"""
from __future__ import annotations

import copy
import itertools
import math
from typing import Dict, List, Tuple

from .sim import ResidentMesh, SimConfig, tune_struct

_cache: Dict[Tuple, dict] = {}

BLOCKS = (256,)
# 0 = the default register budget (4 workgroups per CU; 3 for the Plucker
# kernel); 3 and 5 select the kernels compiled for 3 and 5 per CU
BLOCKS_PER_CU = (0, 3, 5)
CHUNKS = (1, 8)


def sample_batch(sims: List[SimConfig], total: int) -> List[SimConfig]:
    """The batch with each simulation's photon count scaled so that the sum is
    about `total`, keeping at least 256 photons per simulation."""
    n = sum(s.nphoton for s in sims)
    out = []
    for s in sims:
        c = copy.copy(s)
        c.nphoton = max(256, int(math.ceil(s.nphoton * total / max(n, 1))))
        c.issavedet = 0
        out.append(c)
    return out


def autotune(rm: ResidentMesh, sims: List[SimConfig], sample_photons: int = 6_000_000, verbose=False) -> dict:
    if not sims:
        raise ValueError("autotune needs at least one simulation")
    key = (rm.mesh.ne, sims[0].method, min(len(sims), 512) // 16)
    if key in _cache:
        return _cache[key]
    trial = sample_batch(sims, sample_photons)
    best = None
    log = []
    for block, bpc, chunk in itertools.product(BLOCKS, BLOCKS_PER_CU, CHUNKS):
        t = tune_struct(block=block, blocks_per_cu=bpc, chunk=chunk)
        res = rm.run_batch(trial, want_field=False, tune=t)
        steps = sum(r.raytet for r in res)
        rate = steps / res[0].kernel_ms
        log.append({"block": block, "blocks_per_cu": bpc, "chunk": chunk, "steps_per_ms": rate})
        if verbose:
            print(log[-1], flush=True)
        if best is None or rate > best["steps_per_ms"]:
            best = log[-1]
    best = dict(best)
    best["trials"] = log
    _cache[key] = best
    return best


def tune_of(best: dict):
    return tune_struct(block=best["block"], blocks_per_cu=best["blocks_per_cu"], chunk=best["chunk"])
