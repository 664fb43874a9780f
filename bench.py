"""Walk-steps/s of the walk-update hot path on one MI355X.

Headline workload (`ppr4096-rmat26`): batched multi-source PPR on R-MAT
scale 26 (67,108,864 vertices, edge factor 16 = 1,073,741,824 edges, a/b/c/d =
.57/.19/.19/.05, seed 3), 4096 uniform sources x 2000 walks, stop probability
0.15, cap 20, exact sparse per-(source, vertex) u32 counts, one persistent
step kernel per pass.

Secondary workloads (under "workloads"): raw random walks on R-MAT scale 24
(`rawwalks-rmat24`, one walk per vertex, fixed length 10, no restart); SimRank
on the transposed scale-28 graph; continuous-batching restart-PPR serving on
the scale-26 graph in HBM; and serving over a host-resident R-MAT graph of
scale 31, edge factor 24 (51.5 G edges in pinned host memory, a 64 GB
device cache), or the largest scale the host's memory holds.

A bench "step" is one pass of the engine over all walks (gw_run), graph and
walk records resident in HBM before timing.  The launch shape (block threads x
waves per SIMD x walks a wavefront claims at a time) is picked by gw_autotune
before the timed passes, or fixed with --shape; "graph_replay" reports the
same passes replayed from one captured hipGraph.  Prints one JSON line.

roofline.achieved uses the path's algorithmic bytes per walk-step,
B_step = 64 B x (2 + r): one random 64-B sector for the offset pair, one for
the column, and r = 1 more when the workload records a per-step count (the
per-source counts here; raw walks have r = 0), divided by the step kernel's
mean duration (HIP events on the engine stream).  The CPU baseline is the
reference's own engine (an app on its API with msppr's recorder and the same
source list, built from its sources into oracle/_ref/) on the same graph; the
oracle's OpenMP replay is reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import graphwalker_amd as gw  # noqa: E402

PEAK_HBM_GBPS = 8000.0
SECTOR = 64  # bytes per random access at HBM granularity

WORKLOADS = {
    "rawwalks-rmat24": dict(scale=24, ef=16, seed=2, kind=gw.JUMP, L=10, p=0.0, recorder=gw.REC_NONE,
                            block_vertices=1 << 20, sources=0, wps=0,
                            desc="R-MAT scale 24 ef 16 seed 2, 1 walk/vertex, L=10, jump p=0"),
    "rawwalks-rmat20": dict(scale=20, ef=16, seed=2, kind=gw.JUMP, L=10, p=0.0, recorder=gw.REC_NONE,
                            block_vertices=1 << 16, sources=0, wps=0,
                            desc="R-MAT scale 20 ef 16 seed 2, 1 walk/vertex, L=10, jump p=0 (reduced)"),
    # PPR's output is the per-source counts (msppr keeps no walk records either): final walk records
    # are not written (GW_REC_NO_FINAL), for the plain-kernel comparison as well
    "ppr4096-rmat26": dict(scale=26, ef=16, seed=3, kind=gw.STOP, L=20, p=0.15, recorder=gw.REC_SOURCE_COUNT,
                           flags=gw.REC_NO_FINAL, block_vertices=1 << 22, sources=4096, wps=2000,
                           desc="R-MAT scale 26 ef 16 seed 3, 4096 sources x 2000 walks, stop 0.15, cap 20, "
                                "sparse per-source counts (no final walk records)"),
    "ppr256-rmat22": dict(scale=22, ef=16, seed=3, kind=gw.STOP, L=20, p=0.15, recorder=gw.REC_SOURCE_COUNT,
                          flags=gw.REC_NO_FINAL, block_vertices=1 << 20, sources=256, wps=2000,
                          desc="R-MAT scale 22 ef 16 seed 3, 256 sources x 2000 walks (reduced)"),
}


def bytes_per_step(recorder: int) -> int:
    """Algorithmic bytes per walk-step, B_step = 64 x (2 + r): the offset
    pair's sector, the column's sector, and one more (r = 1) when the walk
    program records a per-step count."""
    r = 1 if recorder & (gw.REC_VERTEX_COUNT | gw.REC_SOURCE_COUNT) else 0
    return SECTOR * (2 + r)


def algorithmic_bytes(steps: int, recorder: int) -> int:
    return steps * bytes_per_step(recorder)


def csrc_digest() -> str:
    """sha256 over the library's sources: ties a committed counter profile to
    the kernels it was measured on."""
    import hashlib

    h = hashlib.sha256()
    for d in (os.path.join(ROOT, "graphwalker_amd", "csrc"), os.path.join(ROOT, "include")):
        for f in sorted(os.listdir(d)):
            if f.endswith((".hip", ".h")):
                h.update(f.encode())
                with open(os.path.join(d, f), "rb") as fh:
                    h.update(fh.read())
    return h.hexdigest()


def setup(eng: gw.Engine, wl: dict) -> int:
    n = 1 << wl["scale"]
    t0 = time.perf_counter()
    eng.graph_rmat(wl["scale"], n * wl["ef"], wl["seed"])
    if wl["sources"]:
        src = gw.sources(wl["seed"], n, wl["sources"])
        eng.walks_init_sources(src, wl["wps"], source_field=np.arange(wl["sources"], dtype=np.uint32))
    else:
        eng.walks_init_per_vertex(1)
    eng.graph_blocks(np.arange(0, n + 1, wl["block_vertices"], dtype=np.uint32))
    print(f"[bench] graph + walks ready in {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    return n


def timed(eng: gw.Engine, wl: dict, n: int, sched: str, layout: str, K: int, W: int, source_method=gw.SRC_SORT,
          shape=None):
    prog = gw.program(wl["kind"], wl["L"], wl["p"], n, seed=wl["seed"], recorder=wl["recorder"] | wl.get("flags", 0),
                      schedule=sched, layout=layout, source_method=source_method)
    if shape:
        prog.block_threads, prog.waves_per_simd = shape[0], shape[1]
        prog.walk_chunk = shape[2] if len(shape) > 2 else 0
    for _ in range(W):
        eng.run(prog)
    t = time.perf_counter()
    kms, steps, visits, st = [], 0, 0, None
    for _ in range(K):
        st = eng.run(prog)
        kms.append(st["step_ms"])
        steps += st["steps"]
        visits += st["visits"]
    wall = time.perf_counter() - t
    return prog, wall, steps, visits, kms, st


def replayed(eng: gw.Engine, prog: gw.Program, K: int, W: int = 1) -> dict:
    """K passes of prog replayed from one captured hipGraph (one launch call
    per pass instead of the run's host sequence), after W warm-up replays;
    every replay ends synchronised."""
    eng.run_capture(prog)
    for _ in range(max(1, W)):
        eng.run_replay()
    t = time.perf_counter()
    steps, gms = 0, []
    for _ in range(K):
        st = eng.run_replay()
        steps += st["steps"]
        gms.append(st["total_ms"])
    wall = time.perf_counter() - t
    return {"walk_steps_per_s": steps / wall, "graph_ms": float(np.mean(gms)), "wall_s": wall}


def pool_measure(eng: gw.Engine, wl: dict, n: int, prog: gw.Program, window: int, K: int,
                 keep_final: bool) -> dict:
    """The headline walks as a host walk pool (pinned memory, one record per
    vertex, id == position) stepped `window` walks at a time."""
    walks = eng.host_empty((n,), gw.WALK_DTYPE)
    v = np.arange(n, dtype=np.uint32)
    walks["id"], walks["source"], walks["vertex"], walks["hop"] = v, v, v, 0
    eng.walks_load_host(walks, window, keep_final=keep_final)
    eng.run(prog)
    t = time.perf_counter()
    steps, st = 0, None
    for _ in range(K):
        st = eng.run(prog)
        steps += st["steps"]
    wall = time.perf_counter() - t
    eng.walks_init_per_vertex(1)  # device walks again
    return {"walk_steps_per_s": steps / wall, "window": window, "windows": st["pool_windows"],
            "pcie_bytes_per_pass": st["pool_bytes"], "pass_ms": wall / K * 1e3, "keep_final": keep_final,
            "pcie_GBps": st["pool_bytes"] / (wall / K) / 1e9}


def cpu_baseline(eng: gw.Engine, wl: dict, prog: gw.Program, target_s: float, steps_per_visit: float) -> dict:
    import oracle as O

    off, col = eng.graph_read()
    op = O.Program(prog.kind, prog.L, prog.stop_thr, prog.nvertices, prog.stream, prog.seed)
    n = off.shape[0] - 1
    threads = O.num_threads()
    if wl["sources"]:
        src = gw.sources(wl["seed"], n, wl["sources"])
        starts = np.repeat(src, wl["wps"])
        total = starts.shape[0]
    else:
        starts = None
        total = n
    nw = 1 << 18
    while True:
        v = starts[:nw] if starts is not None else np.arange(nw, dtype=np.uint32) % n
        walks = O.make_walks(v)
        t0 = time.perf_counter()
        r = O.replay(op, off, col, walks, counts=True)
        elapsed = time.perf_counter() - t0
        if elapsed >= target_s / 4 or nw >= total:
            break
        nw = min(total, int(nw * max(2.0, target_s / max(elapsed, 1e-3))))
    port = {"value": r["steps"] / elapsed, "unit": "walk-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle OpenMP replay (per-vertex counts) of the first {nw} walks of the same program on the "
                      f"same device-generated graph: {r['steps']} steps in {elapsed:.2f} s"}
    if wl["sources"]:
        if O.reference_app("msppr_sources") is None:
            return port
        return cpu_reference_multisource(eng, wl, off, col, threads, target_s, steps_per_visit, port)
    if wl["kind"] != gw.JUMP or O.reference_app("rawrandomwalks") is None:
        return port
    # The reference engine itself (its rawrandomwalks app built from its own
    # sources): the same graph in its on-disk layout, as many walks as the
    # workload from uniform sources, L jump steps, its own 0.15 teleport.
    import shutil
    import tempfile

    tmp = tempfile.mkdtemp(prefix="gw_ref_")
    try:
        base = os.path.join(tmp, "graph")
        O.write_reference_graph(base, off, col)
        blocksize_kb = max(1, -(-int(off[-1]) * 4 // 16 // 1024))  # ~16 blocks, as the 2^20-vertex blocks
        R = total
        t0 = time.perf_counter()
        ref = O.run_reference_rawrandomwalks(base, n, R, wl["L"], threads, blocksize_kb, 16, timeout=300)
        wall = time.perf_counter() - t0
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    steps = R * wl["L"]
    return {"value": steps / ref["00_runtime"], "unit": "walk-steps/s", "cores": threads, "kind": "reference",
            "sample": f"reference rawrandomwalks (built from its sources, -O3 -fopenmp, execthreads {threads}) on the "
                      f"same graph in its on-disk layout, {blocksize_kb} KB blocks all in memory: {R} walks x "
                      f"{wl['L']} steps (uniform sources, its 0.15 teleport) in {ref['00_runtime']:.2f} s of its "
                      f"00_runtime ({wall:.1f} s wall)",
            "exec_phase_walk_steps_per_s": steps / ref["5_exec_updates"], "port": port}


def cpu_reference_multisource(eng: gw.Engine, wl: dict, off: np.ndarray, col: np.ndarray, threads: int,
                              target_s: float, steps_per_visit: float, port: dict) -> dict:
    """The reference engine on the batched PPR workload: its RandomWalkwithStop
    program and msppr's recorder (oracle/ref_msppr_sources.cpp) on the same
    graph in its on-disk layout, with the workload's own sources (a leading
    share of them, grown until the run takes about target_s), one in-memory
    block, execthreads = the host's cores; timed by its 00_runtime."""
    import shutil
    import tempfile

    import oracle as O

    n = off.shape[0] - 1
    src = gw.sources(wl["seed"], n, wl["sources"])
    tmp = tempfile.mkdtemp(prefix="gw_ref_")
    try:
        base = os.path.join(tmp, "graph")
        O.write_reference_graph(base, off, col)
        blocksize_kb = -(-int(off[-1]) * 4 // 1024) + 1  # the whole column file in one block
        k = max(1, wl["sources"] // 16)
        while True:
            t0 = time.perf_counter()
            ref = O.run_reference_msppr_sources(base, src[:k], wl["wps"], wl["L"], threads, timeout=600,
                                                blocksize_kb=blocksize_kb)
            wall = time.perf_counter() - t0
            if ref["00_runtime"] >= target_s / 3 or k >= wl["sources"]:
                break
            k = min(wl["sources"], int(k * max(2.0, target_s / max(ref["00_runtime"], 1e-3))))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    visits_per_s = ref["visits"] / ref["00_runtime"]
    return {"value": visits_per_s * steps_per_visit, "unit": "walk-steps/s", "cores": threads, "kind": "reference",
            "sample": f"reference engine (built from its sources, -O3 -fopenmp, execthreads {threads}) running an app "
                      f"on its API with msppr's recorder (a DiscreteDistribution per source and thread) on the same "
                      f"graph in its on-disk layout, one in-memory block: the workload's first {k} of "
                      f"{wl['sources']} sources x {wl['wps']} walks, stop 0.15, cap {wl['L']}: {ref['visits']} visits "
                      f"in {ref['00_runtime']:.2f} s of its 00_runtime ({wall:.1f} s wall); walk-steps = visits x "
                      f"{steps_per_visit:.4f}, the device pass's exact steps per visit",
            "visits_per_s": visits_per_s, "exec_phase_visits_per_s": ref["visits"] / ref["5_exec_updates"],
            "port": port}


def frontend_measure(eng: gw.Engine, wl: dict, n: int, timeout: float = 300) -> dict:
    """The reference's rawrandomwalks app, compiled unchanged against this
    package's GraphWalker API headers (graphwalker_amd/bin/compat/), on the
    graph resident in `eng` written in the on-disk layout: R = n walks of L
    steps from rand() % N sources (its fixed 0.15 teleport), one block, with
    the front end's recorder `none` (the app's updateInfo is empty) and with
    the event log replayed to updateInfo.  Time from its own 00_runtime (walk
    gathering, upload, the device run, the replay)."""
    import re
    import shutil
    import subprocess
    import tempfile

    app = os.path.join(ROOT, "graphwalker_amd", "bin", "compat", "rawrandomwalks")
    if not os.access(app, os.X_OK):
        return {}
    off, col = eng.graph_read()
    tmp = tempfile.mkdtemp(prefix="gw_fe_")
    out = {}
    try:
        base = os.path.join(tmp, "graph")
        kb = -(-int(off[-1]) * 4 // 1024) + 1  # one block
        gw.write_graph_files(base, off, col, kb)
        del off, col
        os.makedirs(os.path.join(tmp, "conf"), exist_ok=True)
        with open(os.path.join(tmp, "conf", "graphwalker.cnf"), "w") as f:
            f.write("# options are given on the command line\n")
        for rec in ("none", "events"):
            t0 = time.perf_counter()
            r = subprocess.run([app, "file", base, "N", str(n), "R", str(n), "L", str(wl["L"]), "blocksize_kb", str(kb),
                                "recorder", rec], cwd=tmp, capture_output=True, text=True, timeout=timeout)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                out[rec] = {"error": f"exit {r.returncode}: {r.stderr[-500:]}"}
                continue
            def timer(key):
                m = re.search(r"\[" + key + r"\] time=([0-9.eE+-]+)s", r.stdout)
                return float(m.group(1)) if m else None
            steps = int(re.search(r"\[walk_steps\] (\d+)", r.stdout).group(1))
            rt = timer("00_runtime")
            out[rec] = {"walk_steps": steps, "runtime_s": rt, "walk_steps_per_s": steps / rt if rt else None,
                        "gather_s": timer("3_getCurrentWalks"), "upload_s": timer("2_findSubGraph"),
                        "device_run_s": timer("5_exec_updates"), "replay_s": timer("z_replay_updateInfo"),
                        "process_wall_s": wall}
            print(f"[bench] front end (unchanged rawrandomwalks, recorder {rec}): {steps / rt / 1e9:.3f} G steps/s "
                  f"({rt:.2f} s of its 00_runtime)", file=sys.stderr, flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out["desc"] = (f"the reference's apps/rawrandomwalks.cpp compiled unchanged against graphwalker_amd/cxx, on the "
                   f"{wl['desc'].split(',')[0]} graph in the on-disk layout (one block): {n} walks x {wl['L']} steps; "
                   f"00_runtime = gathering the packed walk records, their upload, the device run and the replay")
    return out


def measure(eng: gw.Engine, name: str, K: int, W: int, sched: str, layout: str, compare: bool,
            shape=None) -> dict:
    wl = WORKLOADS[name]
    n = setup(eng, wl)
    variants = {}
    tuned = None
    if shape:
        tuned = {"block_threads": shape[0], "waves_per_simd": shape[1],
                 "walk_chunk": shape[2] if len(shape) > 2 else 512, "run_ms": None}
    else:
        # launch shape (block size x waves per SIMD) picked by gw_autotune before the timed passes
        base = gw.program(wl["kind"], wl["L"], wl["p"], n, seed=wl["seed"], recorder=wl["recorder"] | wl.get("flags", 0),
                          schedule=sched,
                          layout=layout)
        best, best_ms = eng.autotune(base, reps=2)
        shape = (best.block_threads, best.waves_per_simd, best.walk_chunk or 512)
        tuned = {"block_threads": shape[0], "waves_per_simd": shape[1], "walk_chunk": shape[2], "run_ms": best_ms}
        print(f"[bench] {name} autotune: {shape[0]} threads x {shape[1]} waves/SIMD, {shape[2]} walks per "
              f"wavefront claim, {best_ms:.2f} ms",
              file=sys.stderr, flush=True)
    if compare:
        for s_, l_ in (("plain", "csr"), ("persistent", "csr"), ("persistent", "ext"), ("blocked", "ext")):
            if (s_, l_) == (sched, layout):
                continue
            _, w_, st_, vi_, k_, last = timed(eng, wl, n, s_, l_, max(2, K // 3), 1)
            variants[f"{s_}-{l_}"] = {"walk_steps_per_s": st_ / w_, "kernel_ms": float(np.mean(k_)),
                                      "rounds": last["rounds"], "bucket_ms": last["bucket_ms"]}
            print(f"[bench] {name} {s_}-{l_}: {st_ / w_ / 1e9:.2f} G steps/s", file=sys.stderr, flush=True)
        # the blocked schedule with its block size and launch shape picked by gw_autotune
        bprog = gw.program(wl["kind"], wl["L"], wl["p"], n, seed=wl["seed"], recorder=wl["recorder"] | wl.get("flags", 0),
                           schedule="blocked", layout="ext")
        bbest, _ = eng.autotune(bprog, reps=1)
        nb = eng.graph_info()["nblocks"]
        _, w_, st_, _, k_, last = timed(eng, wl, n, "blocked", "ext", max(2, K // 3), 1,
                                        shape=(bbest.block_threads, bbest.waves_per_simd, bbest.walk_chunk))
        variants["blocked-ext-tuned"] = {"walk_steps_per_s": st_ / w_, "kernel_ms": float(np.mean(k_)),
                                         "rounds": last["rounds"], "bucket_ms": last["bucket_ms"], "blocks": nb,
                                         "block_vertices": -(-n // nb),
                                         "launch_shape": [bbest.block_threads, bbest.waves_per_simd,
                                                          bbest.walk_chunk or 512]}
        print(f"[bench] {name} blocked-ext tuned ({nb} blocks): {st_ / w_ / 1e9:.2f} G steps/s", file=sys.stderr,
              flush=True)
        eng.graph_blocks(np.arange(0, n + 1, wl["block_vertices"], dtype=np.uint32))
    prog, wall, steps, visits, kms, st = timed(eng, wl, n, sched, layout, K, W, shape=shape)
    per_pass_steps, per_pass_visits = steps // K, visits // K
    kavg = float(np.mean(kms))
    abytes = algorithmic_bytes(per_pass_steps, wl["recorder"])
    achieved = abytes / (kavg * 1e-3) / 1e9
    print(f"[bench] {name} {sched}-{layout}: {steps / wall / 1e9:.2f} G steps/s", file=sys.stderr, flush=True)
    res = {
        "value": steps / wall, "wall_s": wall, "steps": per_pass_steps, "visits": per_pass_visits, "prog": prog,
        "kernel_ms": kavg, "kernel_ms_median": float(np.median(kms)), "algorithmic_bytes": abytes, "achieved": achieved, "layout": st["layout"], "n": n,
        "variants": variants, "wl": wl, "last": st, "tuned": tuned,
    }
    if sched != "blocked":
        r = replayed(eng, prog, K, W)
        res["graph_replay"] = r
        print(f"[bench] {name} {sched}-{layout} graph replay: {r['walk_steps_per_s'] / 1e9:.2f} G steps/s",
              file=sys.stderr, flush=True)
    if "plain-csr" in variants:
        res["speedup_vs_plain"] = res["value"] / variants["plain-csr"]["walk_steps_per_s"]
    res["reduce_ms"] = st["reduce_ms"]
    return res


def serve_sim(eng: gw.Engine, n: int, qps: float, seconds: float, seed: int = 5, hops: int = 10,
              walks_lo: int = 1000, walks_hi: int = 100_000, topk: int = 100, warm_seconds: float = 0.0,
              pool_walks: int = 1 << 24, arena_bytes: int = 16 << 30) -> dict:
    """Poisson arrivals of restart-PPR queries (p = 0.15, 10 hops) at `qps`,
    walk counts log-uniform in [walks_lo, walks_hi], uniform sources, served
    by continuous batching; each query returns its top-k vertices (0 = every
    pair).  Latency = completion (result on the host) minus arrival, on the
    host clock.  One warm-up query is served before the clock starts; with
    warm_seconds, the queries arriving in that first stretch of the stream
    are served but left out of the figures (a host-resident graph's page
    cache adapts to the demand meanwhile)."""
    rng = np.random.default_rng(seed)
    nq = max(1, int(qps * (seconds + warm_seconds)))
    arrivals = np.cumsum(rng.exponential(1.0 / qps, nq))
    walks = np.exp(rng.uniform(np.log(walks_lo), np.log(walks_hi), nq)).astype(np.int64)
    sources = gw.sources(seed, n, nq)
    qarr = np.empty(nq, gw.QUERY_DTYPE)
    qarr["qid"], qarr["source"], qarr["walks"] = np.arange(nq), sources, walks
    eng.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=seed), pool_walks=pool_walks, arena_bytes=arena_bytes,
                   max_queries=4096, hops_per_step=hops)
    if topk:
        eng.serve_topk(topk)
    # one warm-up query served before the clock starts: the serving kernels'
    # first launches (code loading) belong to no query's latency
    eng.serve_submit([(nq, int(sources[0]), walks_lo)])
    for _ in range(1000):
        eng.serve_step()
        r = eng.serve_poll()
        if r is not None:
            eng.serve_result(*r, sort=False)
            break
    lat = np.zeros(nq)
    measured = arrivals >= warm_seconds
    keys = ("steps", "visits", "step_ms", "h2d_bytes", "pages_loaded", "sampled_reads", "sampled_resident")
    acc = {k: 0 for k in keys}  # over the serve steps after warm_seconds
    nsteps = 0
    resident = 0
    done = nxt = 0
    t0 = time.perf_counter()
    t_meas = None
    try:
        while done < nq:
            now = time.perf_counter() - t0
            k = int(np.searchsorted(arrivals, now, side="right"))
            if k > nxt:
                eng.serve_submit(qarr[nxt:k])
                nxt = k
            if done < nxt:
                in_window = now >= warm_seconds
                if in_window and t_meas is None:
                    t_meas = now
                st = eng.serve_step()
                t_done = time.perf_counter() - t0  # the step's completed results are on the host
                if in_window:
                    for k_ in keys:
                        acc[k_] += st[k_]
                    nsteps += 1
                    resident = st["pages_resident"]
                if topk:
                    # every completed query's top-k in one call
                    while (q := eng.serve_results(4096, copy=False)[0]).shape[0]:
                        qi = q.astype(np.int64)
                        lat[qi] = t_done - arrivals[qi]
                        done += qi.shape[0]
                else:
                    while (r := eng.serve_poll()) is not None:
                        eng.serve_result(*r, sort=False)
                        lat[r[0]] = t_done - arrivals[r[0]]
                        done += 1
            elif nxt < nq:
                time.sleep(max(0.0, min(0.001, arrivals[nxt] - now)))
    finally:
        eng.serve_close()
    wall = time.perf_counter() - t0 - (t_meas if t_meas is not None else 0.0)
    lm = lat[measured] if measured.any() else lat
    out = {"qps_offered": qps, "queries": int(measured.sum()), "seconds": wall,
           "p50_ms": float(np.percentile(lm, 50) * 1e3), "p99_ms": float(np.percentile(lm, 99) * 1e3),
           "mean_ms": float(lm.mean() * 1e3), "walk_steps_per_s": acc["steps"] / wall, "visits": acc["visits"],
           "serve_steps": nsteps, "gpu_busy_frac": acc["step_ms"] * 1e-3 / wall,
           # walk-steps per second of serve-step kernel time (HIP events around the kernel)
           "in_kernel_walk_steps_per_s": acc["steps"] / (acc["step_ms"] * 1e-3) if acc["step_ms"] else 0.0,
           "hops_per_step": hops,
           "walks_per_query": f"log-uniform [{walks_lo}, {walks_hi}]", "result": f"top-{topk}" if topk else "all pairs"}
    if warm_seconds:
        out["warmup_queries"] = int((~measured).sum())
        out["warmup_seconds"] = warm_seconds
    if acc["sampled_reads"]:
        out["page_cache"] = {
            "pages_resident": resident, "pages_loaded": acc["pages_loaded"], "h2d_bytes": acc["h2d_bytes"],
            "h2d_GBps": acc["h2d_bytes"] / wall / 1e9,
            "column_reads_from_cache_frac": acc["sampled_resident"] / acc["sampled_reads"],
            # 1-in-64 sample of the column reads, the rest crossed the host link as 4-byte reads
            "host_column_reads_per_s": (acc["sampled_reads"] - acc["sampled_resident"]) * 64 / wall,
        }
    return out


def serve_ooc_measure(scale_max: int, ef: int, seed: int, cache_frac_of_graph: float, cache_cap: int,
                      qps: float, seconds: float, warm_seconds: float, hops: int, topk: int) -> dict:
    """Serving over a host-resident graph: R-MAT (edge factor `ef`, seed) built
    range by range into pinned host memory at the largest scale <= scale_max
    whose CSR fits the host's available memory with room to spare, served
    through a device page cache of cache_cap bytes (offsets included) -- or,
    below scale_max, the same fraction of the graph.  Its own engine, so
    earlier workloads' host memory is released first."""
    avail = 0
    with open("/proc/meminfo") as f:
        for line in f:
            if line.startswith("MemAvailable:"):
                avail = int(line.split()[1]) * 1024
    total = 0
    with open("/proc/meminfo") as f:
        for line in f:
            if line.startswith("MemTotal:"):
                total = int(line.split()[1]) * 1024
    room = min(int(avail * 0.85), 240 << 30)  # the CSR plus the process, under the host's limit
    eng = gw.Engine(0)
    tried = []
    try:
        # the largest scale whose CSR the host can pin (pinned memory may be
        # capped below the available memory: a refused allocation steps down)
        for scale in range(scale_max, 19, -1):
            n, m = 1 << scale, ef << scale
            graph_bytes = (n + 1) * 8 + m * 4
            if graph_bytes > room:
                tried.append(f"scale {scale}: {graph_bytes / 2**30:.0f} GiB of CSR above the available host memory")
                continue
            cache = cache_cap if scale == scale_max else int(graph_bytes * cache_frac_of_graph)
            try:
                t0 = time.perf_counter()
                build_ms = eng.graph_rmat_host(scale, m, seed, cache_bytes=cache)
                build_s = time.perf_counter() - t0
                break
            except gw.GwError as exc:
                if "pinned host memory" not in str(exc):
                    raise
                tried.append(f"scale {scale}: pinning {graph_bytes / 2**30:.0f} GiB refused ({exc})")
                print(f"[bench] {tried[-1]}", file=sys.stderr, flush=True)
        else:
            raise gw.GwError("no scale's CSR could be pinned: " + "; ".join(tried))
        print(f"[bench] host-resident R-MAT scale {scale} ef {ef}: {graph_bytes / 2**30:.1f} GiB built in "
              f"{build_s:.1f} s", file=sys.stderr, flush=True)
        probe_gathers = 1 << 24
        host_read_ms = eng.host_gather_probe(4 << 30, probe_gathers)
        h2d_ms = eng.h2d_probe(1 << 30)
        t0 = time.perf_counter()
        r = serve_sim(eng, n, qps, seconds, seed=seed + 1, hops=hops, topk=topk, warm_seconds=warm_seconds,
                      pool_walks=1 << 25, arena_bytes=32 << 30)
        serve_wall = time.perf_counter() - t0
    finally:
        eng.close()
    host_ceiling = probe_gathers / (host_read_ms * 1e-3)
    h2d_ceiling = (1 << 30) / (h2d_ms * 1e-3) / 1e9
    pc = r.get("page_cache", {})
    r.update({
        "desc": f"R-MAT scale {scale} ef {ef} seed {seed}: {n} vertices, {m} edges, {graph_bytes / 1e9:.1f} GB of CSR in "
                f"pinned host memory; offsets copied to HBM, columns through a {cache / 1e9:.1f} GB device page cache "
                f"(offsets included, pages chosen by sampled demand); restart-PPR queries (p 0.15, 10 hops)",
        "scale": scale, "scale_asked": scale_max, "larger_scales": tried, "edge_factor": ef, "graph_bytes": graph_bytes,
        "cache_bytes": cache,
        "host_mem_total": total, "host_mem_available": avail,
        "graph_build_s": build_s, "graph_build_device_ms": build_ms, "serve_wall_s": serve_wall,
        "host_read_ceiling_per_s": host_ceiling, "h2d_ceiling_GBps": h2d_ceiling,
    })
    if pc:
        pc["h2d_frac_of_link"] = pc["h2d_GBps"] / h2d_ceiling
        pc["host_reads_frac_of_ceiling"] = pc["host_column_reads_per_s"] / host_ceiling
    return r


def stream_measure(eng: gw.Engine, scale: int, seed: int, cache_bytes: int, queries: int, K: int,
                   walks_lo: int = 1000, walks_hi: int = 100_000) -> dict:
    """Out-of-core run on the device-built graph left resident by the last
    workload: the same restart-PPR batch (p = 0.15, 10 hops, per-query
    source counts in the hash table) first on the graph in HBM, then with the
    graph moved to pinned host memory behind a device block cache of
    cache_bytes.  Walks per query log-uniform in [walks_lo, walks_hi], uniform
    sources."""
    n = 1 << scale
    rng = np.random.default_rng(seed)
    src = gw.sources(seed + 11, n, queries)
    wpq = np.exp(rng.uniform(np.log(walks_lo), np.log(walks_hi), queries)).astype(np.int64)
    starts = np.repeat(src, wpq)
    walks = np.zeros(starts.shape[0], dtype=gw.WALK_DTYPE)
    walks["id"] = np.arange(starts.shape[0], dtype=np.uint32)
    walks["source"] = starts
    walks["vertex"] = starts
    lg = int(np.ceil(np.log2(max(1, starts.shape[0]) * 10 * 2)))  # table for ~10 visits per walk at <= 1/2 load

    def batch(sched: str, layout: str) -> tuple[float, dict]:
        prog = gw.program(gw.RESTART, 10, 0.15, n, seed=seed, recorder=gw.REC_SOURCE_COUNT, schedule=sched,
                          layout=layout, source_method=gw.SRC_HASH, table_log2=lg)
        eng.walks_load(walks)
        eng.run(prog)
        t = time.perf_counter()
        st = None
        steps = 0
        for _ in range(K):
            st = eng.run(prog)
            steps += st["steps"]
        return steps / (time.perf_counter() - t), st

    hbm_rate, hbm_st = batch("persistent", "ext")
    hoff = eng.host_empty((eng.n + 1,), np.uint64)
    hcol = eng.host_empty((eng.m,), np.uint32)
    eng.graph_read(hoff, hcol)
    graph_bytes = hoff.nbytes + hcol.nbytes
    eng.graph_set_host(hoff, hcol, cache_bytes)
    rate, st = batch("blocked", "csr")
    # the other load-pass block choices (results are the same; the loads differ)
    policies = {}
    for name, pol, prob in (("min_step_0.2", gw.BLOCKS_MIN_STEP, 0.2), ("max_weight", gw.BLOCKS_MAX_WEIGHT, 0.0)):
        eng.stream_policy(pol, prob, seed=seed)
        r_, st_ = batch("blocked", "csr")
        policies[name] = {"walk_steps_per_s": r_, "blocks_loaded": st_["blocks_loaded"], "rounds": st_["rounds"],
                          "h2d_bytes_per_pass": st_["h2d_bytes"], "same_steps": st_["steps"] == st["steps"]}
        print(f"[bench] stream policy {name}: {r_ / 1e9:.3f} G steps/s, {st_['blocks_loaded']} loads",
              file=sys.stderr, flush=True)
    eng.stream_policy(gw.BLOCKS_MAX_WALKS)
    h2d_ms = eng.h2d_probe(1 << 30)
    out = {
        "desc": f"R-MAT scale {scale} graph ({graph_bytes / 2**30:.2f} GiB CSR) in pinned host memory, device block "
                f"cache {cache_bytes / 2**30:.2f} GiB; {queries} restart-PPR queries (p 0.15, 10 hops), walks per "
                f"query log-uniform [{walks_lo}, {walks_hi}], hash source counts",
        "walks": int(starts.shape[0]), "walk_steps_per_pass": st["steps"],
        "walk_steps_per_s": rate, "in_hbm_walk_steps_per_s": hbm_rate,
        "same_steps_as_in_hbm": st["steps"] == hbm_st["steps"] and st["visits"] == hbm_st["visits"],
        "rounds": st["rounds"], "launches": st["launches"], "blocks": eng.graph_info()["nblocks"], "cache_slots": st["cache_slots"],
        "blocks_loaded": st["blocks_loaded"], "h2d_bytes_per_pass": st["h2d_bytes"],
        "copy_ms": st["copy_ms"], "step_ms": st["step_ms"], "bucket_ms": st["bucket_ms"], "total_ms": st["total_ms"],
        "h2d_GBps_achieved": st["h2d_bytes"] / max(1e-9, st["copy_ms"] * 1e-3) / 1e9,
        "h2d_GBps_probe": (1 << 30) / (h2d_ms * 1e-3) / 1e9,
        "graph_bytes": graph_bytes, "cache_bytes": cache_bytes, "policy": "max_walks", "other_policies": policies,
    }
    print(f"[bench] stream: {rate / 1e9:.3f} G steps/s ({hbm_rate / 1e9:.2f} in HBM), {st['rounds']} rounds, "
          f"{st['blocks_loaded']} loads, {st['h2d_bytes'] / 2**30:.1f} GiB copied at "
          f"{out['h2d_GBps_achieved']:.1f} GB/s (probe {out['h2d_GBps_probe']:.1f})", file=sys.stderr, flush=True)
    return out


def simrank_measure(eng: gw.Engine, scale: int, seed: int, npairs: int, R: int, L: int, c: float, K: int) -> dict:
    """Config 4: SimRank meeting counts on a transposed R-MAT graph (edge
    factor 16): npairs uniform pairs, R walks per endpoint with the
    restart-with-joint program (p 0.15, period L), first-meeting counts per
    offset joined on the device; SimRank = sum_l counts[l] c^l / R^2."""
    n = 1 << scale
    t0 = time.perf_counter()
    eng.graph_rmat(scale, n * 16, seed)
    eng.graph_transpose()
    build_s = time.perf_counter() - t0
    ab = gw.sources(seed, n, 2 * npairs).reshape(npairs, 2)
    a, b = ab[:, 0].copy(), ab[:, 1].copy()
    prog = gw.program(gw.RESTART_JOINT, L, 0.15, n, seed=seed, recorder=gw.REC_EVENTS, event_cap=L - 1,
                      schedule="persistent", layout="ext")
    eng.walks_init_pairs(a, b, R, L)
    eng.run(prog)
    eng.meeting_counts(npairs, R, L)
    steps = 0
    run_ms = []
    t = time.perf_counter()
    for _ in range(K):
        st = eng.run(prog)
        counts = eng.meeting_counts(npairs, R, L)
        steps += st["steps"]
        run_ms.append(st["total_ms"])
    wall = time.perf_counter() - t
    sim = (counts.astype(np.float64) * (c ** np.arange(L))).sum(axis=1) / (R * R)
    out = {
        "desc": f"R-MAT scale {scale} ef 16 seed {seed}, transposed; {npairs} uniform pairs x {R} walks per endpoint, "
                f"restart-with-joint p 0.15, L {L}, c {c}; first-meeting counts joined on the device",
        "walk_steps_per_pass": st["steps"], "walk_steps_per_s": steps / wall, "pairs_per_s": npairs * K / wall,
        "step_ms": float(np.mean(run_ms)), "pass_ms": wall / K * 1e3,
        "join_ms": wall / K * 1e3 - float(np.mean(run_ms)), "graph_build_s": build_s,
        "simrank_mean": float(sim.mean()), "simrank_max": float(sim.max()),
        "pairs_meeting": int((counts.sum(axis=1) > 0).sum()),
    }
    print(f"[bench] simrank: {out['walk_steps_per_s'] / 1e9:.2f} G steps/s, {out['pairs_per_s']:.0f} pairs/s "
          f"({out['pass_ms']:.1f} ms per pass, graph built in {build_s:.1f} s)", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ppr4096-rmat26", choices=sorted(WORKLOADS))
    ap.add_argument("--secondary", default="rawwalks-rmat24", help="extra workload reported under 'workloads' ('' = none)")
    ap.add_argument("--schedule", default="persistent", choices=["plain", "persistent", "blocked"])
    ap.add_argument("--layout", default="ext", choices=["csr", "ext"])
    ap.add_argument("--shape", default="", help="launch shape 'THREADS,WAVES[,CHUNK]' (block threads, waves per "
                                               "SIMD, walks a wavefront claims at a time, default 512); "
                                               "default: picked by the autotuner")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-qps", type=float, nargs="*", default=[2000.0, 20000.0, 40000.0])
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-topk", type=int, default=100, help="vertices returned per query (0 = every pair)")
    ap.add_argument("--serve-hops", type=int, default=10,
                    help="hops per walk per serve step (10 = the program's L: a walk ends within one step)")
    ap.add_argument("--stream-cache-mib", type=int, default=1024,
                    help="out-of-core run: device block cache for the headline's graph (0 = skip)")
    ap.add_argument("--stream-queries", type=int, default=256)
    ap.add_argument("--pool-window", type=int, default=1 << 21,
                    help="walks per window of the host-walk-pool figure on the raw-walk workload (0 = skip)")
    ap.add_argument("--simrank-scale", type=int, default=28, help="SimRank line on R-MAT of this scale (0 = skip)")
    ap.add_argument("--frontend", type=int, default=1, help="time the unchanged rawrandomwalks app (1) or not (0)")
    ap.add_argument("--ooc-scale", type=int, default=31,
                    help="serving over a host-resident R-MAT of this scale, edge factor 24 (or the largest the host "
                         "memory holds; 0 = skip)")
    ap.add_argument("--ooc-cache-gb", type=float, default=64.0, help="its device page cache, offsets included")
    ap.add_argument("--ooc-qps", type=float, default=2000.0)
    ap.add_argument("--ooc-seconds", type=float, default=6.0)
    ap.add_argument("--ooc-warm-seconds", type=float, default=4.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    eng = gw.Engine(0)
    info = eng.device_info()
    print(f"[bench] {info}", file=sys.stderr, flush=True)
    shape = tuple(int(x) for x in args.shape.split(",")) if args.shape else None

    # ---- secondary workloads first (the headline's graph then stays for serving)
    workloads = {}
    pool = {}
    post = {}
    if args.secondary and args.secondary != args.workload:
        s = measure(eng, args.secondary, max(3, args.steps // 2), 1, args.schedule, args.layout, compare=False,
                    shape=shape)
        _, w_, st_, _, _, _ = timed(eng, s["wl"], s["n"], "plain", "csr", 2, 1)
        workloads[args.secondary] = {
            "desc": s["wl"]["desc"], "walk_steps_per_s": s["value"], "walk_steps_per_pass": s["steps"],
            "kernel_ms": s["kernel_ms"], "roofline_achieved_GBps": s["achieved"],
            "roofline_frac": s["achieved"] / PEAK_HBM_GBPS, "bytes_per_step": bytes_per_step(s["wl"]["recorder"]),
            "plain_csr_walk_steps_per_s": st_ / w_, "speedup_vs_plain": s["value"] / (st_ / w_),
            "schedule": f"{args.schedule}-{args.layout}", "launch_shape": s["tuned"], "graph_replay": s.get("graph_replay"),
        }
        # post-processing on this graph: top-k of the per-vertex counts (the toplist caller)
        wl2 = s["wl"]
        eng.run(gw.program(wl2["kind"], wl2["L"], wl2["p"], s["n"], seed=wl2["seed"], recorder=gw.REC_VERTEX_COUNT,
                           schedule=args.schedule, layout=args.layout))
        for k in (100, 1 << 20):
            eng.counts_topk(k)
            t = time.perf_counter()
            eng.counts_topk(k)
            post[f"counts_topk_{k}_ms"] = (time.perf_counter() - t) * 1e3
        print(f"[bench] counts top-k: {post}", file=sys.stderr, flush=True)
        # the same workload through the C++ front end: the reference's app, unchanged
        if not wl2["sources"] and args.frontend:
            workloads[args.secondary]["frontend"] = frontend_measure(eng, wl2, s["n"])
        # the same walks as a host walk pool (out-of-HBM walk pools), stepped in windows
        if args.pool_window and not wl2["sources"]:
            for keep in (False, True):
                r = pool_measure(eng, wl2, s["n"], s["prog"], args.pool_window, max(2, args.steps // 3), keep)
                pool["with_finals" if keep else "counts_only"] = r
                print(f"[bench] host walk pool (window {args.pool_window}, finals {keep}): "
                      f"{r['walk_steps_per_s'] / 1e9:.2f} G steps/s, {r['pcie_GBps']:.1f} GB/s over PCIe",
                      file=sys.stderr, flush=True)

    # ---- the headline
    m = measure(eng, args.workload, args.steps, args.warmup, args.schedule, args.layout, compare=True, shape=shape)
    wl = m["wl"]
    plain_global = plain_hash = global_sort = None
    if wl["sources"]:
        _, w_, st_, _, _, _ = timed(eng, wl, m["n"], "plain", "csr", 2, 1, source_method=gw.SRC_GLOBAL)
        plain_global = st_ / w_
        _, w_, st_, _, _, _ = timed(eng, wl, m["n"], "plain", "csr", 2, 1, source_method=gw.SRC_HASH)
        plain_hash = st_ / w_
        _, w_, st_, _, _, lg_ = timed(eng, wl, m["n"], args.schedule, args.layout, 2, 1, source_method=gw.SRC_GLOBAL)
        global_sort = {"walk_steps_per_s": st_ / w_, "reduce_ms": lg_["reduce_ms"]}
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(eng, wl, m["prog"], args.cpu_seconds, m["steps"] / max(1, m["visits"]))

    # random-gather ceiling: independent 4-byte gathers over 4 GiB
    gathers = 1 << 30
    gms = eng.gather_probe(4 << 30, gathers)
    rand_gbps = gathers * SECTOR / (gms * 1e-3) / 1e9
    # streaming copy of 1 GiB: sequential ceiling, and the known-size kernel
    # that calibrates FETCH_SIZE / WRITE_SIZE in profiles
    copy_bytes = 1 << 30
    cms = eng.copy_probe(copy_bytes)

    # HBM bytes per launch of the step kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # (profiles/), valid only for the library sources they were measured on
    traffic = traffic_copy = None
    traffic_note = "no counter profile for this workload"
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{args.schedule}_{args.layout}.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        if t.get("csrc_sha256") == csrc_digest():
            traffic, traffic_copy = t.get("bytes_per_launch"), t.get("bytes_per_launch_copy_calibrated")
            traffic_note = f"{os.path.relpath(tfile, ROOT)} (same library sources)"
        else:
            traffic_note = f"{os.path.relpath(tfile, ROOT)} was measured on other library sources: not used"

    serving = {}
    if args.serve_qps:
        # continuous batching on the headline's graph, resident in HBM
        for q in args.serve_qps:
            r = serve_sim(eng, m["n"], q, args.serve_seconds, hops=args.serve_hops, topk=args.serve_topk)
            serving[f"qps_{int(q)}"] = r
            print(f"[bench] serving {q:.0f} qps: p50 {r['p50_ms']:.2f} ms p99 {r['p99_ms']:.2f} ms, "
                  f"{r['walk_steps_per_s'] / 1e9:.2f} G steps/s", file=sys.stderr, flush=True)

    stream = {}
    if args.stream_cache_mib:
        stream = stream_measure(eng, wl["scale"], wl["seed"], args.stream_cache_mib << 20, args.stream_queries, 2)

    simrank = {}
    if args.simrank_scale:
        simrank = simrank_measure(eng, args.simrank_scale, 4, 1024, 1000, 11, 0.6, 3)
    eng.close()

    serve_ooc = {}
    if args.ooc_scale:
        try:
            serve_ooc = serve_ooc_measure(args.ooc_scale, 24, 5, 64.0 / 223.0, int(args.ooc_cache_gb * 1e9),
                                          args.ooc_qps, args.ooc_seconds, args.ooc_warm_seconds, args.serve_hops,
                                          args.serve_topk)
            print(f"[bench] serving over host-resident scale {serve_ooc['scale']}: p50 {serve_ooc['p50_ms']:.1f} ms "
                  f"p99 {serve_ooc['p99_ms']:.1f} ms, {serve_ooc['walk_steps_per_s'] / 1e9:.3f} G steps/s",
                  file=sys.stderr, flush=True)
        except gw.GwError as exc:
            serve_ooc = {"error": str(exc)}
            print(f"[bench] serving over a host-resident graph failed: {exc}", file=sys.stderr, flush=True)

    frac = m["achieved"] / PEAK_HBM_GBPS
    roofline = {"bound": "hbm", "achieved": m["achieved"], "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": frac,
                "traffic": traffic, "traffic_copy_calibrated": traffic_copy, "traffic_source": traffic_note,
                "kernel": f"step_{args.schedule}_kernel ({args.layout})", "kernel_ms": m["kernel_ms"],
                "kernel_ms_median": m["kernel_ms_median"], "algorithmic_bytes_per_launch": m["algorithmic_bytes"],
                "bytes_per_step": bytes_per_step(wl["recorder"]),
                "pass_frac": m["algorithmic_bytes"] / ((m["kernel_ms"] + m["reduce_ms"]) * 1e-3) / 1e9 / PEAK_HBM_GBPS}
    if traffic:
        roofline["traffic_frac"] = traffic / (m["kernel_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBPS
    # The headline pass is the captured run replayed (one hipGraph launch a
    # pass, every pass's counters checked on the host); the eager run -- the
    # host enqueueing each kernel -- is reported beside it.
    rep = m.get("graph_replay")
    value, wall = (rep["walk_steps_per_s"], rep["wall_s"]) if rep else (m["value"], m["wall_s"])
    out = {
        "metric": "walk-steps/s",
        "value": value,
        "unit": "walk-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (R-MAT generated on the device)",
        "config": {"workload": args.workload, "desc": wl["desc"], "walk_steps_per_pass": m["steps"],
                   "visits_per_pass": m["visits"], "schedule": args.schedule, "layout": args.layout,
                   "launch": "graph replay" if rep else "eager", "device": info["name"]},
        "eager": {"walk_steps_per_s": m["value"], "ms_per_step": m["wall_s"] / args.steps * 1e3},
        "roofline": roofline,
        "random_access_ceiling": {"gathers_per_s": gathers / (gms * 1e-3), "GBps_at_64B": rand_gbps,
                                  "frac_of_ceiling": m["achieved"] / rand_gbps},
        "copy_GBps": 2 * copy_bytes / (cms * 1e-3) / 1e9,
        "variants": m["variants"],
        "launch_shape": m["tuned"],
        "graph_replay": m.get("graph_replay"),
        "workloads": workloads,
        "serving": serving,
        "serving_out_of_core": serve_ooc,
        "out_of_core": stream,
        "workloads_simrank": simrank,
        "post_processing": post,
        "host_walk_pool": pool,
    }
    if "speedup_vs_plain" in m:
        out["speedup_vs_plain_same_recorder"] = m["speedup_vs_plain"]
    if plain_global:
        out["speedup_vs_plain_global_sort"] = m["value"] / plain_global
        out["speedup_vs_plain_hash"] = m["value"] / plain_hash
        out["plain_recorders_walk_steps_per_s"] = {"global_sort": plain_global, "hash": plain_hash}
        out["global_sort"] = global_sort
        out["source_method"] = {gw.SRC_PARTITIONED: "partitioned", gw.SRC_GLOBAL: "global-sort",
                                gw.SRC_HASH: "hash"}.get(m["last"]["source_method"], "?")
        out["reduce_ms"] = m["reduce_ms"]
        out["flushed_sources"] = m["last"]["flushed_sources"]
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
