"""mmc_amd -- MI355X-native mesh Monte Carlo photon transport.

The transport runs in hand-written HIP kernels for gfx950 inside
``mmc_amd/lib/libmmc_hip.so`` (C ABI: include/mmc_hip.h).  This package is
the Python host layer: mesh preparation, simulation settings, plain and
batched runs, and the serving loop.  There is no CPU fallback.
"""
from . import meshgen
from .mesh import TetMesh
from .sim import ResidentMesh, SimConfig, SimResult, normalize_field, photon_seeds, run

__all__ = ["TetMesh", "SimConfig", "SimResult", "ResidentMesh", "run", "normalize_field",
           "photon_seeds", "meshgen"]
