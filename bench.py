#!/usr/bin/env python3
"""Benchmark of the MI355X photon tracer (one JSON line on stdout).

Default workload (the plain drop-in path): the 100x100x50 mm layered slab with
a spherical inclusion, 3.0M tetrahedra, 1e8 photons per step, 50 time gates
of 0.1 ns, element fluence, 16 disk detectors.  A step is one full simulation
on the device-resident mesh: the persistent transport kernel plus the FP64
final reduction of the fluence field.  The mesh upload and the host copy of
the field are outside the timed region.

    python bench.py [--steps K] [--warmup W] [--photons N] [--mode plain|batch|all]

The JSON line carries `roofline` for the transport kernel (algorithmic bytes
per ray-tetrahedron step x steps per launch / kernel time from HIP events on
the launch stream) and `cpu_baseline` (the CPU oracle, a restatement of the
reference's OpenMP path, on this host's cores over a bounded photon sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBPS = 8000.0
BYTES_PER_STEP_ELEM = 92.0   # 64 B face planes + 16 B face neighbours + 4 B label + 8 B FP32 atomic RMW
# Plucker nodal step: 48 B node coordinates + 16 B node ids + 16 B neighbours
# + 4 x 8 B nodal FP32 atomic RMW + 16 B next-element ids
BYTES_PER_STEP_NODAL = 128.0
# the same accesses counted in whole 128-byte cache lines, each access in
# its own line: Badouel planes, neighbours, the deposit's line; Plucker node
# coordinates (48 B may straddle two lines), node ids, neighbours, four nodal
# deposits
LINE_BYTES_PER_STEP_ELEM = 3 * 128.0
LINE_BYTES_PER_STEP_NODAL = 8 * 128.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(mesh_arrays, sim, budget_s=15.0, threads=None, nodal=False):
    """Time the CPU oracle (OpenMP restatement of mmc_run_mp, built -O3
    -msse4.1 like the reference's sse target) on a bounded photon sample of
    the same workload, with every core this process may run on."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle as O
    node, elem, et, med = mesh_arrays[:4]
    om = O.OracleMesh(node, elem, et, med, nodal=nodal, srcelem=mesh_arrays[4] if len(mesh_arrays) > 4 else None)
    avail = len(os.sched_getaffinity(0))
    threads = threads or avail
    import copy
    s = copy.copy(sim)
    done, ms = 0, 0.0
    chunk = 2000 * threads
    t0 = time.time()
    while time.time() - t0 < budget_s:
        s.nphoton = chunk
        s.seed = sim.seed + done
        r = O.run(om, s, nthreads=threads)
        done += chunk
        ms += r.runtime_ms
    return {"value": done / ms, "unit": "photons/ms", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(), "affinity_cpus": avail,
            "sample": f"{done} photons of the same workload (glibc-libm oracle, {threads} OpenMP threads on "
                      f"{avail} available of {os.cpu_count()} host CPUs, {ms / 1000:.1f} s)",
            "calibration": "none: the reference itself cannot be compiled here (its mmc_utils.c needs the "
                           "generated mmc_core.clh), so no reference/oracle speed ratio exists"}


def bench_plain(args):
    """The plain drop-in path on a device-resident mesh: the slab workload
    (default) or, for profiling the Plucker nodal kernel, --workload cube
    (the 60 mm cube of 2 mm cells) or --workload shells (the 24.0M-tet
    jittered shells mesh, one 64x64 pattern request, 100 gates)."""
    import mmc_amd as M
    from mmc_amd import workloads as W
    from mmc_amd import meshgen
    t = time.time()
    cube = args.workload == "cube"
    shells = args.workload == "shells"
    nodal = cube or shells
    if shells:
        arrays = W.shells_arrays(1.0)
        mesh = M.TetMesh(*arrays[:4], srcelem=arrays[4]).prepare(nodal=True)
    else:
        arrays = meshgen.cube60(2.0) if cube else meshgen.slab_layers(1.0)
        mesh = M.TetMesh(*arrays).prepare(nodal=cube)
    res = M.ResidentMesh(mesh)
    log(f"{args.workload} mesh: {mesh.ne} tets, prep+upload {time.time() - t:.1f} s")
    if shells:
        sim = W.shells_sims(8, photons=8 * args.photons)[0]
        sim.e0 = int(mesh.srcelem[0])
    else:
        sim = W.cube60_sim(args.photons) if cube else W.slab_sim(args.photons)
    bytes_per_step = BYTES_PER_STEP_NODAL if nodal else BYTES_PER_STEP_ELEM
    for i in range(args.warmup):
        r = res.run_batch([sim], want_field=False)[0]
        log(f"warmup {i}: {r.kernel_ms:.1f} ms kernel, {r.total_ms:.1f} ms total")
    steps_ms, kern_ms, raytet, photons = [], [], 0.0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        sim.photon_offset = (i + args.warmup) * args.photons
        ts = time.perf_counter()
        r = res.run_batch([sim], want_field=False)[0]
        steps_ms.append((time.perf_counter() - ts) * 1e3)
        kern_ms.append(r.kernel_ms)
        raytet += r.raytet
        photons += args.photons
        log(f"step {i}: {steps_ms[-1]:.1f} ms, kernel {r.kernel_ms:.1f} ms, {r.raytet / args.photons:.0f} steps/photon, "
            f"absorbed {r.absorbweight / r.launchweight:.4f}, detected {r.detected}")
    total_ms = (time.perf_counter() - t0) * 1e3
    value = photons / total_ms
    k_avg = float(np.mean(kern_ms))
    steps_per_launch = raytet / args.steps
    achieved = steps_per_launch * bytes_per_step / (k_avg * 1e-3) / 1e9
    # counters of the same kernel and workload (rocprofv3 --pmc summary in profiles/)
    counters = None
    cfile = os.path.join(HERE, "profiles", f"transport_{args.workload}_counters.json")
    if os.path.exists(cfile):
        with open(cfile) as f:
            counters = json.load(f)
    traffic = counters["hbm_bytes_per_step"] * steps_per_launch if counters else None
    kname = "mmc::transport<1, false, 3, 1>" if nodal else "mmc::transport<0, false, 4, 1>"
    if shells:
        workload = (f"shells: jittered 200x200x100 mm box, 1 mm cells x6 = {mesh.ne / 1e6:.1f}M tets, 5 tissues in "
                    f"nested ellipsoidal shells, one 64x64 pattern over 60x60 mm, {args.photons:.0e} photons/step, "
                    "100 gates x 0.05 ns, nodal fluence (Plucker), 32 detectors")
    elif cube:
        workload = (f"cube: 60 mm, 2 mm cells x6 = {mesh.ne} tets, mua 0.005, mus 1.0, g 0.01, n 1.37, pencil at "
                    f"(30,30,0), {args.photons:.0e} photons/step, 1 gate 0-5 ns, nodal fluence (Plucker)")
    else:
        workload = ("slab: 100x100x50 mm, 1 mm cells x6 = 3.0M tets, 4 layers + 10 mm sphere, "
                    f"pencil at (50,50,0), {args.photons:.0e} photons/step, 50 gates x 0.1 ns, "
                    "element fluence (branch-less Badouel), 16 disk detectors")
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
            "kernel": kname, "kernel_ms": k_avg,
            "bytes_per_step": bytes_per_step,
            "line_bytes_per_step": LINE_BYTES_PER_STEP_NODAL if nodal else LINE_BYTES_PER_STEP_ELEM,
            "steps_per_launch": steps_per_launch}
    if counters:
        # traffic: the HBM bytes per step that FETCH_SIZE/WRITE_SIZE measured
        # for this kernel and workload (the committed counter summary) times
        # this run's steps per launch
        roof["traffic_source"] = (f"{counters['hbm_bytes_per_step']:.1f} measured HBM bytes per step "
                                  f"({counters['source']}) x steps per launch of this run")
        # the kernel issues vector instructions nearly every cycle: the
        # HBM fraction above is the algorithmic-bytes view, the limiter is
        # VALU issue (figures from the committed counter summary)
        roof["limiter"] = "valu-issue" if counters.get("valu_issue_frac", 0) > 0.8 else "latency/divergence"
        roof["counters"] = {k: counters[k] for k in ("source", "valu_issue_frac", "active_lane_frac",
                                                    "valu_per_wave_step", "l2_hit_frac", "hbm_bytes_per_step")
                            if k in counters}
    out = {
        "metric": "photons/ms on 1 GPU (plain vs batched/served) + p50/p99 solve latency",
        "value": value, "unit": "photons/ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload,
                   "photons_per_step": args.photons, "tets": int(mesh.ne), "gates": sim.gates()[0],
                   "steps_per_photon": raytet / photons, "ray_tet_steps_per_ms": raytet / sum(kern_ms)},
        "roofline": roof,
    }
    res.release()
    return out, arrays, sim


def bench_batch(args):
    """The batched workload on the slab mesh: one persistent launch for a set
    of simulations with their own pencil source, media and photon count,
    against the plain drop-in (upload, run, free) once per simulation."""
    import mmc_amd as M
    from mmc_amd import autotune as AT
    from mmc_amd import meshgen
    from mmc_amd import workloads as W
    mesh = M.TetMesh(*meshgen.slab_layers(1.0)).prepare(nodal=False)
    sims = W.slab_batch_sims(256)[: args.batch_sims]
    photons = sum(s.nphoton for s in sims)
    rm = M.ResidentMesh(mesh)
    t = time.perf_counter()
    best = AT.autotune(rm, sims)
    tune_s = time.perf_counter() - t
    log(f"autotune ({tune_s:.1f} s): block {best['block']}, blocks/CU {best['blocks_per_cu']}, chunk {best['chunk']}")
    t = time.perf_counter()
    res = rm.run_batch(sims, want_field=False, tune=AT.tune_of(best))
    batch_ms = (time.perf_counter() - t) * 1e3
    steps = sum(r.raytet for r in res)
    log(f"batched: {len(sims)} simulations, {photons} photons, {batch_ms:.0f} ms, {steps / photons:.0f} steps/photon")
    rm.release()
    t = time.perf_counter()
    plain_steps = 0.0
    for s in sims:
        r = M.run(mesh, s, want_field=False)
        plain_steps += r.raytet
    plain_ms = (time.perf_counter() - t) * 1e3
    log(f"plain drop-in, one call per simulation: {plain_ms:.0f} ms")
    return {"workload": f"slab mesh (3.0M tets), first {len(sims)} of the 256-simulation batch: pencils on a "
                        "16x16 grid of 4 mm pitch, per-tissue media from seed 7, photon counts log-uniform "
                        "in [1e5, 1e7], 10 gates x 0.5 ns, element fluence",
            "simulations": len(sims), "photons": photons,
            "batched_photons_per_ms": photons / batch_ms, "plain_photons_per_ms": photons / plain_ms,
            "speedup": plain_ms / batch_ms, "same_steps": plain_steps == steps,
            "autotune": {k: best[k] for k in ("block", "blocks_per_cu", "chunk")}, "autotune_s": tune_s}


def bench_serve(args, deadline):
    """The serving workload: Poisson arrivals (200 requests/s) over three
    resident meshes (60 mm cube 162k tets, slab 3.0M tets, jittered box
    5.2M tets), served by the continuous-batching loop and, for comparison,
    a sample of the same requests by the plain drop-in one request at a
    time.  The stream is cut to the requests whose photons fit the time left
    before `deadline` (at the plain slab rate), which the line states."""
    import mmc_amd as M
    from mmc_amd import meshgen
    from mmc_amd import serve as S
    t = time.time()
    meshes = {"cube": M.TetMesh(*meshgen.cube60(2.0)).prepare(nodal=False),
              "slab": M.TetMesh(*meshgen.slab_layers(1.0)).prepare(nodal=False),
              "jitter": M.TetMesh(*meshgen.jittered_box(120, 120, 60, 1.0, amp=0.2)).prepare(nodal=False)}
    log(f"serving meshes ready in {time.time() - t:.1f} s: " + ", ".join(f"{k} {m.ne}" for k, m in meshes.items()))
    hi = args.serve_max_photons
    asked = args.serve_requests
    reqs = S.poisson_requests(meshes, asked, rate=200.0, photons=(1e5, hi))
    offered = sum(r.cfg.nphoton for r in reqs) / (reqs[-1].arrival * 1e3)      # photons/ms asked for
    # photons that fit: served once, plus the plain sample; 0.7 of the plain slab rate
    rate = 0.7 * args.plain_rate
    left_ms = max(0.0, (deadline - time.time()) * 1e3)
    every = max(1, args.serve_plain_every)
    n, cum = 0, 0.0
    for r in reqs:
        extra = r.cfg.nphoton * (1.0 + (1.0 / every))
        if (cum + extra) / rate > left_ms:
            break
        cum += extra
        n += 1
    n = max(n, min(8, asked))
    reqs = reqs[:n]
    log(f"serving {n} of {asked} requests ({sum(r.cfg.nphoton for r in reqs):.3g} photons; offered load "
        f"{offered:.3g} photons/ms)")
    srv = S.Server(meshes, arena_bytes=16 << 30, policy=args.serve_policy)
    served = srv.serve(reqs)
    srv.close()
    log(f"served: {served}")
    sample = S.poisson_requests(meshes, asked, rate=200.0, photons=(1e5, hi))[:n][::every]
    plain = S.serve_plain(meshes, sample)
    log(f"plain ({len(sample)} requests): {plain}")
    return {"workload": f"Poisson 200 requests/s, meshes cube 162k / slab 3.0M / jittered 5.2M tets, photons "
                        f"log-uniform [1e5, {hi:.0e}], 1-50 gates, per-tissue media as the batch workload, "
                        "element fluence",
            "requests_asked": asked, "requests": n, "policy": args.serve_policy,
            "offered_photons_per_ms": offered, "capacity_photons_per_ms": served["photons_per_ms"],
            "served": served,
            "plain_sample": {"every": every, **plain},
            "speedup": served["photons_per_ms"] / plain["photons_per_ms"]}


def bench_large(args):
    """The large-scale workload: the 24.0M-tet jittered shells mesh, a 60x60 mm
    planar source with 8 64x64 Bernoulli patterns, 100 gates of 0.05 ns,
    Plucker nodal fluence with FP64 final reduction, 32 detectors,
    args.large_photons photons over the 8 patterns.  Run twice with the same
    statistics per pattern: as 8 pattern requests of large_photons/8 photons
    through the serving loop (the stated configuration), and as one run of
    large_photons/8 photons carrying all 8 patterns (srcnum = 8, 8 slots per
    node and gate) in launches of args.large_slice."""
    import mmc_amd as M
    from mmc_amd import serve as S
    from mmc_amd import workloads as W
    t = time.time()
    mesh = W.shells_mesh(1.0)
    log(f"shells mesh: {mesh.ne} tets, {mesh.nn} nodes, {len(mesh.srcelem)} launch candidates, "
        f"built in {time.time() - t:.1f} s")
    total = int(args.large_photons)
    sl = int(args.large_slice)
    e0 = int(mesh.srcelem[0])

    # 8 pattern requests through the continuous-batching loop
    sims = W.shells_sims(8, photons=total)
    for c in sims:
        c.e0 = e0
    reqs = [S.Request(rid=i, mesh="shells", cfg=c, arrival=0.0) for i, c in enumerate(sims)]
    srv = S.Server({"shells": mesh}, arena_bytes=16 << 30, slice_photons=sl)
    st = srv.serve(reqs)
    srv.close()
    del srv
    log(f"large, 8 pattern requests served: {st}")

    # the same patterns sharing every photon
    rm = M.ResidentMesh(mesh)
    log(f"mesh resident: {rm.device_bytes / 2**30:.2f} GiB")
    shared_n = total // 8
    photons = steps = absorbed = launched = kern = 0.0
    detected = 0
    t = time.perf_counter()
    done = 0
    while done < shared_n:
        n = min(sl, shared_n - done)
        cfg = W.shells_shared_sim(shared_n)
        cfg.nphoton, cfg.photon_offset, cfg.e0 = n, done, e0
        r = rm.run_batch([cfg], want_field=False)[0]
        done += n
        photons += n
        steps += r.raytet
        launched += r.launchweight
        absorbed += r.absorbweight
        kern += r.kernel_ms
        detected += r.detected
        log(f"large, shared patterns: {done:.3g}/{shared_n:.3g} photons, launch {r.kernel_ms:.0f} ms")
    wall = (time.perf_counter() - t) * 1e3
    rm.release()
    return {"workload": f"jittered 200x200x100 mm box, 1 mm cells x6 = {mesh.ne / 1e6:.1f}M tets, 5 tissues in "
                        "nested ellipsoidal shells, 60x60 mm planar source with 8 64x64 Bernoulli(0.5) patterns, "
                        f"{total:.0e} photons over the 8 patterns ({total // 8:.3g} per pattern), 100 gates x "
                        "0.05 ns, Plucker nodal fluence (FP32 atomics, FP64 reduction), 32 detectors",
            "tets": int(mesh.ne), "nodes": int(mesh.nn),
            "served_8_requests": {"photons": st["photons"], "wall_ms": st["wall_ms"],
                                  "pattern_photons_per_ms": st["photons"] / st["wall_ms"],
                                  "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "launches": st["launches"],
                                  "kernel_ms": st["kernel_ms"]},
            "shared_srcnum8": {"photons": int(photons), "launches": -(-shared_n // sl), "wall_ms": wall,
                               "kernel_ms": kern, "photons_per_ms": photons / wall,
                               "pattern_photons_per_ms": 8 * photons / wall,
                               "steps_per_photon": steps / photons, "absorbed_fraction": absorbed / launched,
                               "detected": detected},
            "shared_over_served": (8 * photons / wall) / (st["photons"] / st["wall_ms"])}


def copy_probe(nbytes: int):
    """Device-to-device copy of a known size through the HIP runtime: the
    calibration run for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=int, default=100_000_000)
    ap.add_argument("--workload", default="slab", choices=["slab", "cube", "shells"],
                    help="plain-path workload: the slab (default), the 60 mm cube or the 24.0M-tet shells "
                         "mesh with one pattern request (both Plucker nodal)")
    ap.add_argument("--mode", default="all", choices=["plain", "batch", "serve", "all", "large"],
                    help="all (default): the plain slab steps, then the 256-simulation batch and the "
                         "served request stream")
    ap.add_argument("--large-photons", type=float, default=1e9)
    ap.add_argument("--large-slice", type=float, default=2.5e8)
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-plain-every", type=int, default=5,
                    help="the plain drop-in runs every n-th request of the served stream")
    ap.add_argument("--serve-max-photons", type=float, default=1e8)
    ap.add_argument("--serve-policy", default="srpt", choices=["srpt", "fair"],
                    help="serving order: fewest photons left first (srpt) or first come with equal "
                         "shares (fair)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-sims", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--time-budget", type=float, default=480.0,
                    help="seconds for the whole run; the served stream is cut to fit")
    ap.add_argument("--copy-probe", type=int, default=0, help="only copy this many bytes device-to-device")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): photons shard across GPUs "
              "with disjoint seed ranges and one reduction of the fluence, which is not run here",
              file=sys.stderr)
        sys.exit(2)
    deadline = time.time() + args.time_budget - (0 if args.no_cpu else args.cpu_seconds + 10)
    out, arrays, sim = bench_plain(args)
    args.plain_rate = out["value"]
    if args.mode in ("batch", "all"):
        out["batched"] = bench_batch(args)
    if args.mode in ("serve", "all"):
        out["serving"] = bench_serve(args, deadline)
    if args.mode == "large":
        out["large"] = bench_large(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(arrays, sim, budget_s=args.cpu_seconds,
                                           nodal=args.workload in ("cube", "shells"))
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
