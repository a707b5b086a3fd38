"""Host-side logic of the serving loop (no GPU): the field-arena allocator,
HBM admission with a small fake budget, and chunk planning."""
import numpy as np

import mmc_amd as M
from mmc_amd.serve import FieldArena, Request, Scheduler


def test_arena_first_fit_and_coalescing():
    a = FieldArena(1024)
    x = a.alloc(100)
    y = a.alloc(200)
    z = a.alloc(64)
    assert (x, y, z) == (0, 128, 384)
    assert a.alloc(1024) == -1
    a.release(y, 200)
    assert a.alloc(150) == 128          # reuses the hole
    a.release(x, 100)
    a.release(128, 150)
    a.release(z, 64)
    assert a.free == [(0, 1024)] and a.available == 1024


def mkreq(i, mesh, gates, n, t):
    cfg = M.SimConfig(nphoton=n, method="S", tend=5e-9, tstep=5e-9 / gates)
    return Request(rid=i, mesh=mesh, cfg=cfg, arrival=t)


def test_admission_by_field_bytes_and_fifo_order():
    s = Scheduler(arena_floats=10_000, datalen={"a": 1000, "b": 3000}, slice_photons=1000, min_chunk=100)
    reqs = [mkreq(0, "a", 5, 5000, 0.0), mkreq(1, "b", 2, 500, 0.1), mkreq(2, "a", 2, 300, 0.2)]
    for r in reqs:
        s.arrive(r)
    admitted = s.admit(0.3)
    # request 0 needs 5000 floats, request 1 needs 6000: only request 0 fits
    assert [r.rid for r in admitted] == [0]
    assert [r.rid for r in s.waiting] == [1, 2]
    mesh, chunks = s.plan()
    assert mesh == "a" and [(r.rid, n) for r, n in chunks] == [(0, 1000)]
    reqs[0].assigned = reqs[0].done = 5000
    s.complete(reqs[0], 1.0)
    admitted = s.admit(1.0)
    assert [r.rid for r in admitted] == [1, 2]   # FIFO: 1 first, then 2 fits too
    mesh, chunks = s.plan()
    assert mesh == "b" and chunks[0][1] == 500


def test_request_larger_than_arena_is_rejected():
    s = Scheduler(arena_floats=100, datalen={"a": 1000})
    try:
        s.arrive(mkreq(0, "a", 1, 10, 0.0))
    except ValueError:
        return
    raise AssertionError("oversized request admitted")


def test_fair_share_of_the_slice():
    s = Scheduler(arena_floats=10**6, datalen={"a": 10}, slice_photons=1000, min_chunk=100)
    for i, n in enumerate([10_000, 150, 10_000, 10_000]):
        s.arrive(mkreq(i, "a", 1, n, 0.0))
    s.admit(0.0)
    _, chunks = s.plan()
    assert [n for _, n in chunks] == [250, 150, 250, 250]


def test_plan_skips_requests_whose_photons_are_all_in_flight():
    # with launches in flight, planning uses the photons not yet handed out
    s = Scheduler(arena_floats=10**6, datalen={"a": 10, "b": 10}, slice_photons=1000, min_chunk=100)
    r0, r1 = mkreq(0, "a", 1, 800, 0.0), mkreq(1, "b", 1, 5000, 0.1)
    for r in (r0, r1):
        s.arrive(r)
    s.admit(0.2)
    mesh, chunks = s.plan()
    assert mesh == "a" and [(r.rid, n) for r, n in chunks] == [(0, 800)]
    r0.assigned += 800                # launched, not finished
    mesh, chunks = s.plan()
    assert mesh == "b" and [(r.rid, n) for r, n in chunks] == [(1, 1000)]
    assert r0 in s.active             # completes only when its launch finishes


def test_rf_request_reserves_its_imaginary_block():
    """An RF request (omega > 0) writes its imaginary part right after the
    real one, so it reserves twice the floats; the request admitted next to
    it starts after both blocks."""
    s = Scheduler(arena_floats=10_000, datalen={"a": 1000}, slice_photons=1000, min_chunk=100)
    rf = mkreq(0, "a", 2, 500, 0.0)
    rf.cfg.omega = 1e8
    plain = mkreq(1, "a", 2, 500, 0.1)
    s.arrive(rf)
    s.arrive(plain)
    assert rf.flen == 2 * plain.flen == 4000
    s.admit(0.2)
    assert plain.off >= rf.off + rf.flen


def test_srpt_admits_and_runs_the_fewest_photons_first():
    """Shortest remaining work first: admission by photon count (any request
    whose field fits, not only the head of the queue) and launches filled
    with the requests that have the fewest photons left."""
    s = Scheduler(arena_floats=11_000, datalen={"a": 1000, "b": 1000}, slice_photons=1000, min_chunk=100,
                  policy="srpt")
    reqs = [mkreq(0, "a", 5, 5000, 0.0), mkreq(1, "b", 6, 300, 0.1), mkreq(2, "a", 3, 700, 0.2),
            mkreq(3, "a", 1, 200, 0.3)]
    for r in reqs:
        s.arrive(r)
    admitted = s.admit(0.4)
    # fields: 5000, 6000, 3000, 1000 floats (64-float aligned); smallest
    # photon counts first: 3 (200), 1 (300), 2 (700) fit in 11000 floats,
    # 0 (5000 photons) waits
    assert [r.rid for r in admitted] == [3, 1, 2]
    assert [r.rid for r in s.waiting] == [0]
    mesh, chunks = s.plan()
    # request 3 has the fewest photons left: mesh a, then 2 on the same mesh
    assert mesh == "a" and [(r.rid, n) for r, n in chunks] == [(3, 200), (2, 700)]
    for r, n in chunks:
        r.assigned += n
    mesh, chunks = s.plan()
    assert mesh == "b" and [(r.rid, n) for r, n in chunks] == [(1, 300)]


def test_unknown_policy_is_refused():
    try:
        Scheduler(arena_floats=10, datalen={"a": 1}, policy="lifo")
    except ValueError:
        return
    raise AssertionError("unknown policy accepted")


def test_requests_without_photons_are_admitted_but_never_planned():
    """A zero-photon request takes its (empty) field region and is completed
    by the server at admission; no launch ever carries a chunk of it."""
    for policy in ("fair", "srpt"):
        s = Scheduler(arena_floats=10_000, datalen={"a": 100}, slice_photons=1000, min_chunk=100, policy=policy)
        reqs = [mkreq(0, "a", 1, 0, 0.0), mkreq(1, "a", 1, 1500, 0.0)]
        for r in reqs:
            s.arrive(r)
        assert sorted(r.rid for r in s.admit(0.0)) == [0, 1]
        seen = []
        while True:
            mesh, chunks = s.plan()
            if not chunks:
                break
            assert all(n > 0 for _, n in chunks)
            for r, n in chunks:
                r.assigned += n
                seen.append(r.rid)
        assert set(seen) == {1} and reqs[1].assigned == 1500


def test_statistics_of_an_empty_stream():
    from mmc_amd.serve import summarize
    st = summarize([], 0.5, 0, 0.0)
    assert st["requests"] == 0 and st["photons"] == 0 and st["photons_per_ms"] == 0.0
    r = mkreq(0, "a", 1, 0, 0.0)
    r.admitted, r.finished = 0.0, 0.0
    st = summarize([r], 0.0, 0, 0.0)
    assert st["requests"] == 1 and st["photons"] == 0 and st["p50_ms"] == 0.0
