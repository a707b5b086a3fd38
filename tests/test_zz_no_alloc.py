"""The allocation-free entry point (mmc_hip_run_chunks_noalloc) against the
launch it mirrors (mmc_hip_run_chunks).

Both run the same simulations on a resident mesh into arenas created
beforehand; the fields (FP32 atomics, so equal up to the order of the adds),
energy tallies, ray-tetrahedron counts, detected photons and per-photon
records must agree.  Two calls in a row, the second with a smaller input,
show that the shared scratch array carries nothing over from the previous
call; a launch whose buffers exceed the array is refused with a message.
"""
import ctypes as C
import copy

import numpy as np
import pytest

import mmc_amd as M
from mmc_amd import _lib
from mmc_amd._lib import ResultC, SimC, f64p, ptr
from mmc_amd import workloads as W

pytestmark = pytest.mark.gpu

NOALLOC_SCRATCH = 8 << 20


@pytest.fixture(scope="module")
def slab():
    arrays = M.meshgen.slab_layers(2.0)
    tm = M.TetMesh(*arrays).prepare(nodal=False)
    rm = M.ResidentMesh(tm)
    yield tm, rm
    rm.release()


def sims_for(tm, nsim, nphoton, seed):
    cfgs = []
    for i in range(nsim):
        c = W.slab_sim(nphoton + 997 * i, seed=seed + i)
        c.maxdetphoton = 4000
        cfgs.append(c)
    return cfgs


def launch(entry, tm, rm, cfgs, arena, records=(0, 500)):
    L = _lib.lib()
    n = len(cfgs)
    sims = (SimC * n)()
    keep = []
    offs = np.zeros(n, np.uint64)
    off = 0
    bufs = []
    for i, c in enumerate(cfgs):
        st = c.cstruct(tm)
        keep.append(st._keep)
        sims[i] = st
        offs[i] = off
        off += c.datalen(tm) * c.gates()[0]
        bufs.append(M.sim._ResultBuf(tm, c, want_field=False, records=records if i == 0 else None))
    res = (ResultC * n)()
    for i, b in enumerate(bufs):
        res[i] = b.c
    t = M.sim.tune_struct()
    rc = getattr(L, entry)(rm.handle, sims, n, arena, ptr(offs, C.c_uint64), C.byref(t), res)
    if rc != 0:
        return rc, _lib.last_error(), None
    out = []
    for i, c in enumerate(cfgs):
        b = bufs[i]
        b.c = res[i]
        r = b.result()
        flen = c.datalen(tm) * c.gates()[0]
        field = np.zeros(flen, np.float64)
        s = C.c_double()
        _lib.check(L.mmc_hip_arena_collect(arena, C.c_uint64(int(offs[i])), C.c_uint64(flen), ptr(field, C.c_double),
                                           C.byref(s)), "mmc_hip_arena_collect")
        out.append((r, field))
    return 0, "", out


@pytest.fixture(scope="module")
def arenas(slab):
    tm, rm = slab
    L = _lib.lib()
    hs = []
    for _ in range(2):
        h = C.c_void_p()
        _lib.check(L.mmc_hip_arena_create(C.c_uint64(4 * tm.ne * 50), 0, 8, C.byref(h)), "mmc_hip_arena_create")
        hs.append(h)
    yield hs
    for h in hs:
        L.mmc_hip_arena_release(h)


def compare(a, b):
    for (ra, fa), (rb, fb) in zip(a, b):
        assert ra.launchweight == pytest.approx(rb.launchweight, rel=1e-12)
        assert ra.absorbweight == pytest.approx(rb.absorbweight, rel=1e-9)
        assert ra.escapeweight == pytest.approx(rb.escapeweight, rel=1e-9)
        assert ra.raytet == rb.raytet
        assert ra.detected == rb.detected > 0
        ia, ib = np.argsort(ra.detphoton), np.argsort(rb.detphoton)
        assert np.array_equal(ra.detphoton[ia], rb.detphoton[ib])
        assert np.array_equal(ra.detdata[ia], rb.detdata[ib])
        assert fa.sum() > 0
        # the same photons deposit the same values; only the order of the FP32
        # atomic adds differs between the launches
        assert np.allclose(fa, fb, rtol=2e-5, atol=1e-6 * fa.max())
    pa, pb = a[0][0].photons, b[0][0].photons
    assert pa is not None and len(pa) > 0 and np.array_equal(pa, pb)


def test_noalloc_launch_equals_the_allocating_launch_twice(slab, arenas):
    tm, rm = slab
    big = sims_for(tm, 3, 20000, seed=11)
    small = sims_for(tm, 2, 6000, seed=50)
    for cfgs in (big, small):
        print(f"noalloc: {len(cfgs)} simulations of {[c.nphoton for c in cfgs]} photons", flush=True)
        rc0, err0, ref = launch("mmc_hip_run_chunks", tm, rm, cfgs, arenas[0])
        assert rc0 == 0, err0
        rc1, err1, got = launch("mmc_hip_run_chunks_noalloc", tm, rm, cfgs, arenas[1])
        assert rc1 == 0, err1
        compare(got, ref)


def test_noalloc_refuses_buffers_beyond_the_scratch(slab, arenas):
    tm, rm = slab
    cfgs = sims_for(tm, 1, 1000, seed=3)
    c = copy.copy(cfgs[0])
    c.maxdetphoton = NOALLOC_SCRATCH // 4       # detector records alone exceed the array
    rc, err, _ = launch("mmc_hip_run_chunks_noalloc", tm, rm, [c], arenas[1], records=None)
    assert rc != 0 and "scratch" in err
