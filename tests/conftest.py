"""Shared test setup.

`gpu` tests need an MI355X and the built libmmc_hip.so; everything else runs
on the CPU (the oracle, host-side preparation, ABI exports).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


def _ensure_built():
    lib = os.path.join(ROOT, "mmc_amd", "lib", "libmmc_hip.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "mmc_amd", "csrc")], check=True)
    for name in ("liborc.so", "liborc_pm.so"):
        if not os.path.exists(os.path.join(ROOT, "oracle", "_build", name)):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


def read_dat(path, cols, has_id=True):
    """Read a reference-format .dat mesh file: a header line, then rows of
    [id] + cols values (face-neighbour files carry no id column)."""
    import numpy as np
    first = 1 if has_id else 0
    with open(path) as f:
        f.readline()
        rows = [line.split() for line in f if line.strip()]
    return np.array([[float(x) for x in r[first:first + cols]] for r in rows])
