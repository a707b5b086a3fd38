"""Photon trajectories ("-D M", dlTraj): one record (photon id, position,
weight, element) at launch, after every scattering event and at the end of
the photon (savedebugdata, src/mmc_raytrace.c:1739-1756, 1804, 2050, 2109),
checked against the per-photon end records and scattering counts of the same
run, which the oracle pins photon by photon (tests/test_gpu_parity.py)."""
import os
import struct
import subprocess

import numpy as np
import pytest

import pyoracle as O

import mmc_amd as M
from mmc_amd import meshgen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_trajectories_follow_each_photon():
    arrays = meshgen.slab_layers(2.0, nx=40, ny=40, nz=20)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    cfg = M.SimConfig(nphoton=500, seed=3, method="S", srcpos=(20.0, 20.0, 0.0), srcdir=(0.0, 0.0, 1.0),
                      tstart=0.0, tend=5e-9, tstep=5e-9, outputtype="x", trajectory=True, minenergy=0.0)
    g = M.run(tm, cfg, records=(0, cfg.nphoton))
    ph = g.photons
    nscat = ph["nscat"].astype(np.int64)
    assert g.ntraj == len(g.traj) == int((nscat + 2).sum())
    order = np.argsort(g.traj["id"], kind="stable")        # each photon's records stay in time order
    t = g.traj[order]
    start = np.concatenate([[0], np.cumsum(nscat + 2)[:-1]])
    assert np.array_equal(t["id"], np.repeat(np.arange(cfg.nphoton, dtype=np.uint32), nscat + 2))
    first, last = t[start], t[start + nscat + 1]
    assert np.all(first["p"] == np.float32([20.0, 20.0, 0.0])) and np.all(first["w"] == 1.0)
    assert np.array_equal(last["p"], ph["p"]) and np.array_equal(last["w"], ph["w"])
    assert np.array_equal(last["eid"], ph["eid"])
    # without roulette (minenergy 0) the weight never grows along a trajectory
    mid = np.ones(len(t), bool)
    mid[start] = False
    assert np.all(t["w"][mid] <= t["w"][np.flatnonzero(mid) - 1] * (1 + 1e-6))
    # the oracle sees the same number of scattering events per photon
    om = O.OracleMesh(*arrays, nodal=False)
    o = O.run(om, cfg, photonseed=M.photon_seeds(cfg.seed, 0, cfg.nphoton), records=(0, cfg.nphoton),
              math="portable")
    assert np.array_equal(o.photons["nscat"], ph["nscat"])


def test_trajectory_buffer_keeps_the_first_records():
    arrays = meshgen.slab_layers(2.0, nx=40, ny=40, nz=20)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    cfg = M.SimConfig(nphoton=200, seed=4, method="S", srcpos=(20.0, 20.0, 0.0), srcdir=(0.0, 0.0, 1.0),
                      tstart=0.0, tend=5e-9, tstep=5e-9, trajectory=True, maxjumpdebug=100)
    g = M.run(tm, cfg)
    assert len(g.traj) == 100 and g.ntraj > 100


def test_cli_writes_the_trajectory_file(tmp_path):
    cube2 = os.path.join(ROOT, "tests", "golden", "cube2", "cube2.json")
    r = subprocess.run([os.path.join(ROOT, "mmc_amd", "bin", "mmc"), "-f", cube2, "-n", "300", "-M", "S", "-C",
                        "0", "-D", "TPM", "-s", "tr"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    raw = open(tmp_path / "tr.mct", "rb").read()
    head = struct.unpack("4s7If I f I i I 2i", raw[:64])
    assert head[0] == b"MCXH" and head[4] == 6 and head[6] == 0     # colcount, detected
    n = head[7]
    assert n >= 2 * 300 and len(raw) == 64 + 24 * n
    rec = np.frombuffer(raw[64:], np.float32).reshape(n, 6)
    ids = rec[:, 0].view(np.uint32)
    assert set(np.unique(ids)) == set(range(300))
