"""Continuous-batching photon-transport server.

Meshes stay resident on the device.  Requests (a mesh, a source, a media set,
a photon count, a gate count) are admitted when their output field fits in
the device field arena (admission by HBM), and are then run in photon chunks:
every launch of the persistent kernel takes a share of the remaining photons
of every admitted request on one mesh, so requests that arrive while others
are on the GPU join at the next launch and their photons run in the same
launch as everyone else's.  A request completes when its last chunk has run;
its FP32 field is then reduced to FP64 and its arena space is freed.

Because a photon's random stream depends only on (seed, photon id), a request
produces the same photons whether it runs alone, in one launch, or split
across any number of launches in any mix.
"""
from __future__ import annotations

import copy
import ctypes as C
import dataclasses
import threading
import time
from collections import deque
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from ._lib import ResultC, SimC, f64p, ptr
from .mesh import TetMesh
from .sim import ResidentMesh, SimConfig, tune_struct


@dataclasses.dataclass
class Request:
    rid: int
    mesh: str
    cfg: SimConfig
    arrival: float = 0.0              # seconds after the server clock starts
    want_field: bool = False
    # filled by the server
    flen: int = 0
    off: int = -1
    assigned: int = 0                 # photons handed to launches
    done: int = 0                     # photons whose launches have finished
    admitted: float = -1.0
    finished: float = -1.0
    launches: int = 0
    launchweight: float = 0.0
    absorbweight: float = 0.0
    escapeweight: float = 0.0
    raytet: float = 0.0
    detected: int = 0
    fieldsum: float = 0.0
    field: Optional[np.ndarray] = None
    field_im: Optional[np.ndarray] = None      # RF: imaginary part

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class FieldArena:
    """First-fit allocator over a range of floats (the device field arena).
    Pure host bookkeeping; the device buffer is owned by the server."""

    ALIGN = 64

    def __init__(self, nfloat: int):
        self.n = int(nfloat)
        self.free = [(0, self.n)]       # sorted (offset, length)

    def alloc(self, n: int) -> int:
        n = -(-int(n) // self.ALIGN) * self.ALIGN
        for i, (o, l) in enumerate(self.free):
            if l >= n:
                if l == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (o + n, l - n)
                return o
        return -1

    def release(self, off: int, n: int):
        n = -(-int(n) // self.ALIGN) * self.ALIGN
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((o, l))
        self.free = merged

    @property
    def available(self) -> int:
        return sum(l for _, l in self.free)


class Scheduler:
    """Admission and chunk planning (host logic, testable without a GPU)."""

    def __init__(self, arena_floats: int, datalen: Dict[str, int], slice_photons: int = 2_000_000,
                 min_chunk: int = 100_000, nodelen: Optional[Dict[str, int]] = None,
                 meshes: Optional[Dict[str, TetMesh]] = None, policy: str = "fair"):
        """policy "fair": first-come-first-served admission and an equal share
        of each launch for every admitted request on the launch's mesh;
        "srpt": the requests with the fewest photons left are admitted and
        run first (shortest remaining work first), which lowers the median
        latency of a backlog without changing its throughput."""
        if policy not in ("fair", "srpt"):
            raise ValueError(f"unknown scheduling policy {policy!r}")
        self.policy = policy
        self.arena = FieldArena(arena_floats)
        self.datalen = datalen          # elements per mesh (element fields)
        self.nodelen = nodelen or {}    # nodes per mesh (nodal fields)
        self.meshes = meshes or {}      # meshes, for the voxel counts of dual-grid requests
        self.slice_photons = slice_photons
        self.min_chunk = min_chunk
        self.waiting: deque = deque()
        self.active: List[Request] = []

    def rows(self, r: Request) -> int:
        if r.cfg.is_grid:
            return r.cfg.datalen(self.meshes[r.mesh])
        return self.nodelen[r.mesh] if r.cfg.basisorder else self.datalen[r.mesh]

    def field_len(self, r: Request) -> int:
        """Arena floats of a request: rows x gates x patterns, twice that for
        RF (the imaginary block follows the real one)."""
        return self.rows(r) * r.cfg.gates()[0] * r.cfg.srcnum * (2 if r.cfg.is_rf else 1)

    def arrive(self, r: Request, mesh: Optional[TetMesh] = None):
        if mesh is not None and r.cfg.e0 <= 0 and r.cfg.srctype == "pencil":
            r.cfg.e0 = mesh.locate(r.cfg.srcpos)[0]      # located once, reused by every chunk
        r.flen = self.field_len(r)
        if r.flen > self.arena.n:
            raise ValueError(f"request {r.rid} needs {r.flen} floats, arena holds {self.arena.n}")
        self.waiting.append(r)

    def admit(self, now: float) -> List[Request]:
        """Admission while the request's output field fits the arena: in
        arrival order, stopping at the first that does not fit ("fair"), or
        smallest photon count first, admitting every one that fits ("srpt")."""
        out = []
        if self.policy == "srpt":
            for r in sorted(self.waiting, key=lambda q: (q.cfg.nphoton, q.arrival, q.rid)):
                off = self.arena.alloc(r.flen)
                if off < 0:
                    continue
                self.waiting.remove(r)
                r.off, r.admitted = off, now
                self.active.append(r)
                out.append(r)
            return out
        while self.waiting:
            r = self.waiting[0]
            off = self.arena.alloc(r.flen)
            if off < 0:
                break
            self.waiting.popleft()
            r.off, r.admitted = off, now
            self.active.append(r)
            out.append(r)
        return out

    def plan(self):
        """"fair": pick the mesh of the oldest admitted request and give every
        admitted request on that mesh a share of the launch's photon budget.
        "srpt": pick the mesh of the request with the fewest photons left and
        fill the budget with that mesh's requests, fewest photons left first."""
        open_ = self.plannable()
        if not open_:
            return None, []
        if self.policy == "srpt":
            left = lambda r: r.cfg.nphoton - r.assigned
            mesh = min(open_, key=lambda r: (left(r), r.arrival, r.rid)).mesh
            budget = self.slice_photons
            chunks = []
            for r in sorted((r for r in open_ if r.mesh == mesh), key=lambda r: (left(r), r.arrival, r.rid)):
                if budget <= 0:
                    break
                n = min(left(r), max(budget, self.min_chunk))
                chunks.append((r, n))
                budget -= n
            return mesh, chunks
        mesh = min(open_, key=lambda r: (r.arrival, r.rid)).mesh
        group = [r for r in open_ if r.mesh == mesh]
        share = max(self.min_chunk, self.slice_photons // len(group))
        chunks = [(r, min(share, r.cfg.nphoton - r.assigned)) for r in group]
        return mesh, [(r, n) for r, n in chunks if n > 0]

    def plannable(self) -> List[Request]:
        return [r for r in self.active if r.assigned < r.cfg.nphoton]

    def complete(self, r: Request, now: float):
        self.active.remove(r)
        self.arena.release(r.off, r.flen)
        r.finished = now


class Server:
    def __init__(self, meshes: Dict[str, TetMesh], arena_bytes: int = 8 << 30, slice_photons: int = 8_000_000,
                 tune=None, compensated: bool = False, replicas: int = 8, policy: str = "fair"):
        """arena_bytes of FP32 field space for admitted requests, held as
        `replicas` per-XCD copies (device memory = arena_bytes * replicas,
        twice that when compensated); policy: the scheduler's ("fair" or
        "srpt")."""
        self.res = {k: ResidentMesh(m) for k, m in meshes.items()}
        self.meshes = meshes
        L = _lib.lib()
        h = C.c_void_p()
        nfloat = arena_bytes // 4
        _lib.check(L.mmc_hip_arena_create(C.c_uint64(nfloat), int(compensated), int(replicas), C.byref(h)),
                   "mmc_hip_arena_create")
        self.arena = h
        self.sched = Scheduler(nfloat, {k: m.ne for k, m in meshes.items()}, slice_photons,
                               nodelen={k: m.nn for k, m in meshes.items()}, meshes=meshes, policy=policy)
        self.tune = tune if tune is not None else tune_struct()
        self.launches = 0
        self.kernel_ms = 0.0
        # two launch slots: the next launch waits queued behind the running
        # one and its workgroups fill the CUs the running launch's tail frees
        self.slots = []
        for _ in range(2):
            sh = C.c_void_p()
            _lib.check(L.mmc_hip_slot_create(C.byref(sh)), "mmc_hip_slot_create")
            self.slots.append(sh)
        self.inflight = deque()       # (slot, chunks, results, keep-alive)

    def close(self):
        L = _lib.lib()
        while self.inflight:
            self._finish()
        for sh in self.slots:
            L.mmc_hip_slot_release(sh)
        self.slots = []
        for r in self.res.values():
            r.release()
        if self.arena:
            L.mmc_hip_arena_release(self.arena)
            self.arena = None

    def _launch(self, mesh: str, chunks):
        """Enqueue one launch on a free slot (returns at once)."""
        L = _lib.lib()
        busy = {id(x[0]) for x in self.inflight}
        slot = next(sh for sh in self.slots if id(sh) not in busy)
        rm = self.res[mesh]
        n = len(chunks)
        sims = (SimC * n)()
        offs = np.zeros(n, np.uint64)
        keep = [sims, offs]
        for i, (r, cnt) in enumerate(chunks):
            c = copy.copy(r.cfg)
            c.nphoton = cnt
            c.photon_offset = r.cfg.photon_offset + r.assigned
            st = c.cstruct(rm.mesh)
            keep.append(st._keep)
            sims[i] = st
            offs[i] = r.off
        res = (ResultC * n)()
        _lib.check(L.mmc_hip_run_chunks_async(rm.handle, sims, n, self.arena, ptr(offs, C.c_uint64),
                                              C.byref(self.tune), res, slot), "mmc_hip_run_chunks_async")
        for r, cnt in chunks:
            r.assigned += cnt
            r.launches += 1
        self.inflight.append((slot, chunks, res, keep))
        self.launches += 1

    def _finish(self) -> List[Request]:
        """Wait for the oldest launch in flight; returns the requests whose
        last photons it ran."""
        L = _lib.lib()
        slot, chunks, res, keep = self.inflight.popleft()
        _lib.check(L.mmc_hip_slot_wait(slot, res), "mmc_hip_slot_wait")
        finished = []
        for i, (r, cnt) in enumerate(chunks):
            q = res[i]
            r.done += cnt
            r.launchweight += q.launchweight
            r.absorbweight += q.absorbweight
            r.escapeweight += q.escapeweight
            r.raytet += q.raytet
            r.detected += q.detected
            if r.done >= r.cfg.nphoton:
                finished.append(r)
        self.kernel_ms += res[0].kernel_ms
        return finished

    def _collect(self, r: Request):
        L = _lib.lib()
        s = C.c_double()
        shape = (r.cfg.gates()[0], self.sched.rows(r)) + ((r.cfg.srcnum,) if r.cfg.srcnum > 1 else ())
        nre = r.flen // 2 if r.cfg.is_rf else r.flen
        # the real block, then (RF) the imaginary block that follows it
        for part in range(2 if r.cfg.is_rf else 1):
            buf = np.zeros(shape, np.float64) if r.want_field else None
            host = ptr(buf, C.c_double) if r.want_field else C.cast(None, f64p)
            _lib.check(L.mmc_hip_arena_collect(self.arena, C.c_uint64(r.off + part * nre), C.c_uint64(nre), host,
                                               C.byref(s)), "mmc_hip_arena_collect")
            if part == 0:
                r.field, r.fieldsum = buf, s.value
            else:
                r.field_im = buf

    def serve(self, requests: List[Request]) -> dict:
        """Run a recorded arrival stream (arrival times in seconds from now)
        to completion; returns latency and throughput statistics.  Up to two
        launches are in flight; the host plans and enqueues the next launch
        while the current one runs."""
        L = _lib.lib()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        done = []
        while pending or self.sched.waiting or self.sched.active or self.inflight:
            while pending and pending[0].arrival <= now():
                r = pending.popleft()
                self.sched.arrive(r, self.meshes[r.mesh])
            for r in self.sched.admit(now()):
                if r.cfg.nphoton == 0:
                    # nothing to run: its field is its (cleared) arena region
                    self._collect(r)
                    self.sched.complete(r, now())
                    done.append(r)
            if len(self.inflight) < len(self.slots):
                mesh, chunks = self.sched.plan()
                if chunks:
                    self._launch(mesh, chunks)
                    continue
            if not self.inflight:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
                continue
            # both slots busy, or nothing more to plan: take the oldest launch
            # once it is done (polling, so that arrivals are admitted meanwhile)
            if len(self.inflight) < len(self.slots) and pending and L.mmc_hip_slot_query(self.inflight[0][0]) == 0:
                time.sleep(0.0005)
                continue
            for r in self._finish():
                self._collect(r)
                self.sched.complete(r, now())
                done.append(r)
        wall = now()
        return summarize(done, wall, self.launches, self.kernel_ms)


def summarize(done: List[Request], wall_s: float, launches: int, kernel_ms: float) -> dict:
    if not done:
        return {"requests": 0, "photons": 0, "wall_ms": wall_s * 1e3, "photons_per_ms": 0.0, "p50_ms": 0.0,
                "p99_ms": 0.0, "mean_ms": 0.0, "launches": launches, "kernel_ms": kernel_ms}
    lat = np.array([r.latency for r in done]) * 1e3
    photons = sum(r.cfg.nphoton for r in done)
    first = min(r.arrival for r in done)
    return {"requests": len(done), "photons": photons, "wall_ms": wall_s * 1e3,
            "photons_per_ms": photons / (max(wall_s - first, 1e-9) * 1e3),
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "launches": launches, "kernel_ms": kernel_ms}


def serve_plain(meshes: Dict[str, TetMesh], requests: List[Request]) -> dict:
    """The same arrival stream through the plain drop-in: one mmc_hip_run
    (upload, run, free) per request, first come first served."""
    from .sim import run
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0
    done = []
    kernel_ms = 0.0
    while pending:
        r = pending.popleft()
        if r.arrival > now():
            time.sleep(r.arrival - now())
        res = run(meshes[r.mesh], r.cfg, want_field=r.want_field)
        kernel_ms += res.kernel_ms
        r.finished = now()
        r.launchweight, r.absorbweight, r.raytet = res.launchweight, res.absorbweight, res.raytet
        r.fieldsum, r.field = res.fieldsum, res.field
        done.append(r)
    return summarize(done, now(), len(done), kernel_ms)


def poisson_requests(meshes: Dict[str, TetMesh], n: int, rate: float = 200.0, seed: int = 4,
                     photons=(1e5, 1e8), gates=(1, 50), prop_ranges=None) -> List[Request]:
    """The serving workload: Poisson arrivals, a random resident mesh, a pencil
    on its top face tilted up to 30 degrees, per-tissue media (mua U(0.001,
    0.05), mus' U(0.5, 2), g U(0.8, 0.95), n U(1.33, 1.45)), log-uniform photon
    counts and 1-50 gates over 0-5 ns."""
    rng = np.random.default_rng(seed)
    keys = sorted(meshes)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        key = keys[int(rng.integers(len(keys)))]
        m = meshes[key]
        lo, hi = m.node.min(axis=0), m.node.max(axis=0)
        x = rng.uniform(lo[0] + 0.2 * (hi[0] - lo[0]), hi[0] - 0.2 * (hi[0] - lo[0]))
        y = rng.uniform(lo[1] + 0.2 * (hi[1] - lo[1]), hi[1] - 0.2 * (hi[1] - lo[1]))
        th, ph = np.radians(rng.uniform(0, 30)), rng.uniform(0, 2 * np.pi)
        d = (np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th))
        med = np.zeros((m.prop + 1, 4), np.float32)
        med[0] = [0, 0, 1, 1]
        for k in range(1, m.prop + 1):
            g = rng.uniform(0.8, 0.95)
            med[k] = [rng.uniform(0.001, 0.05), rng.uniform(0.5, 2.0) / (1 - g), g, rng.uniform(1.33, 1.45)]
        nph = int(round(10 ** rng.uniform(np.log10(photons[0]), np.log10(photons[1]))))
        ng = int(rng.integers(gates[0], gates[1] + 1))
        cfg = SimConfig(nphoton=nph, seed=50000 + i, method="S", srcpos=(x, y, float(lo[2])), srcdir=d,
                        tstart=0.0, tend=5e-9, tstep=5e-9 / ng, med=med)
        out.append(Request(rid=i, mesh=key, cfg=cfg, arrival=t))
    return out
