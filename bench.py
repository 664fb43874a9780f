"""Walk-steps/s of the walk-update hot path on one MI355X.

Default workload (`rawwalks-rmat24`): raw random walks on R-MAT scale 24
(16,777,216 vertices, edge factor 16 = 268,435,456 edges, a/b/c/d =
.57/.19/.19/.05, seed 2), one walk per vertex, fixed length 10, no restart
(dangling vertices teleport): 167,772,160 walk-steps per pass.

Secondary workload (`ppr4096-rmat26`, reported under "workloads"): batched
multi-source PPR on R-MAT scale 26 (67,108,864 vertices, 1,073,741,824 edges,
seed 3), 4096 uniform sources x 2000 walks, stop probability 0.15, cap 20,
exact sparse per-(source, vertex) u32 counts.

A bench "step" is one pass of the engine over all walks (gw_run), graph and
walk records resident in HBM before timing.  The launch shape (block threads x
waves per SIMD x walks a wavefront claims at a time) is picked by gw_autotune
before the timed passes, or fixed with --shape; "graph_replay" reports the same passes replayed from one
captured hipGraph.  Prints one JSON line.  The CPU
baseline is the reference's own engine (its rawrandomwalks app, built from its
sources into oracle/_ref/) on the same graph and walk count, on the host's
cores; the oracle's OpenMP replay of the same walk program is reported beside
it (and is the baseline where the reference app is not built).
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


def random_sectors_per_step(layout: int) -> int:
    """Random reads per walk-step at 64-B sector granularity: the CSR layout
    reads one sector for the offset pair and one for the column; the
    successor-header layout reads one 16-B edge record."""
    return 2 if layout == gw.LAYOUT_CSR else 1


def algorithmic_bytes(steps: int, visits: int, recorder: int, layout: int) -> int:
    """Random-access bytes one pass needs: the step reads, plus one sector
    read-modify-write per visit for a count recorder (vertex atomics or the
    source-count hash)."""
    rec = 1 if recorder & (gw.REC_VERTEX_COUNT | gw.REC_SOURCE_COUNT) else 0
    return SECTOR * (steps * random_sectors_per_step(layout) + visits * rec)


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


def replayed(eng: gw.Engine, prog: gw.Program, K: int) -> dict:
    """K passes of prog replayed from one captured hipGraph (one launch call
    per pass instead of the run's host sequence)."""
    eng.run_capture(prog)
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


def cpu_baseline(eng: gw.Engine, wl: dict, prog: gw.Program, target_s: float) -> dict:
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
    if wl["sources"] or wl["kind"] != gw.JUMP or O.reference_app("rawrandomwalks") is None:
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
    abytes = algorithmic_bytes(per_pass_steps, per_pass_visits, wl["recorder"], st["layout"])
    achieved = abytes / (kavg * 1e-3) / 1e9
    print(f"[bench] {name} {sched}-{layout}: {steps / wall / 1e9:.2f} G steps/s", file=sys.stderr, flush=True)
    res = {
        "value": steps / wall, "wall_s": wall, "steps": per_pass_steps, "visits": per_pass_visits, "prog": prog,
        "kernel_ms": kavg, "kernel_ms_median": float(np.median(kms)), "algorithmic_bytes": abytes, "achieved": achieved, "layout": st["layout"], "n": n,
        "variants": variants, "wl": wl, "last": st, "tuned": tuned,
    }
    if sched != "blocked":
        r = replayed(eng, prog, K)
        res["graph_replay"] = r
        print(f"[bench] {name} {sched}-{layout} graph replay: {r['walk_steps_per_s'] / 1e9:.2f} G steps/s",
              file=sys.stderr, flush=True)
    if "plain-csr" in variants:
        res["speedup_vs_plain"] = res["value"] / variants["plain-csr"]["walk_steps_per_s"]
    return res


def serve_sim(eng: gw.Engine, n: int, qps: float, seconds: float, seed: int = 5, hops: int = 10,
              walks_lo: int = 1000, walks_hi: int = 100_000, topk: int = 100) -> dict:
    """Poisson arrivals of restart-PPR queries (p = 0.15, 10 hops) at `qps`,
    walk counts log-uniform in [walks_lo, walks_hi], uniform sources, served
    by continuous batching; each query returns its top-k vertices (0 = every
    pair).  Latency = completion (result on the host) minus arrival, on the
    host clock.  One warm-up query is served before the clock starts."""
    rng = np.random.default_rng(seed)
    nq = max(1, int(qps * seconds))
    arrivals = np.cumsum(rng.exponential(1.0 / qps, nq))
    walks = np.exp(rng.uniform(np.log(walks_lo), np.log(walks_hi), nq)).astype(np.int64)
    sources = gw.sources(seed, n, nq)
    eng.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=seed), pool_walks=1 << 24, arena_bytes=16 << 30,
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
    steps = visits = nsteps = 0
    step_ms = 0.0
    done = nxt = 0
    t0 = time.perf_counter()
    try:
        while done < nq:
            now = time.perf_counter() - t0
            k = nxt
            while k < nq and arrivals[k] <= now:
                k += 1
            if k > nxt:
                eng.serve_submit([(i, int(sources[i]), int(walks[i])) for i in range(nxt, k)])
                nxt = k
            if done + 0 < nxt:
                st = eng.serve_step()
                steps += st["steps"]
                visits += st["visits"]
                step_ms += st["step_ms"]
                nsteps += 1
                t_done = time.perf_counter() - t0
                while (r := eng.serve_poll()) is not None:
                    eng.serve_result(*r, sort=False)
                    lat[r[0]] = t_done - arrivals[r[0]]
                    done += 1
            elif nxt < nq:
                time.sleep(max(0.0, min(0.001, arrivals[nxt] - now)))
    finally:
        eng.serve_close()
    wall = time.perf_counter() - t0
    return {"qps_offered": qps, "queries": nq, "seconds": wall, "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "walk_steps_per_s": steps / wall, "visits": visits, "serve_steps": nsteps,
            "gpu_busy_frac": step_ms * 1e-3 / wall, "hops_per_step": hops,
            "walks_per_query": f"log-uniform [{walks_lo}, {walks_hi}]", "result": f"top-{topk}" if topk else "all pairs"}


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
    ap.add_argument("--workload", default="rawwalks-rmat24", choices=sorted(WORKLOADS))
    ap.add_argument("--secondary", default="ppr4096-rmat26", help="extra workload reported under 'workloads' ('' = none)")
    ap.add_argument("--schedule", default="persistent", choices=["plain", "persistent", "blocked"])
    ap.add_argument("--layout", default="ext", choices=["csr", "ext"])
    ap.add_argument("--shape", default="", help="launch shape 'THREADS,WAVES[,CHUNK]' (block threads, waves per "
                                               "SIMD, walks a wavefront claims at a time, default 512); "
                                               "default: picked by the autotuner")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-qps", type=float, nargs="*", default=[2000.0, 20000.0])
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-topk", type=int, default=100, help="vertices returned per query (0 = every pair)")
    ap.add_argument("--serve-hops", type=int, default=10,
                    help="hops per walk per serve step (10 = the program's L: a walk ends within one step)")
    ap.add_argument("--stream-cache-mib", type=int, default=1024,
                    help="out-of-core run: device block cache for the last workload's graph (0 = skip)")
    ap.add_argument("--stream-queries", type=int, default=256)
    ap.add_argument("--pool-window", type=int, default=1 << 21,
                    help="walks per window of the host-walk-pool figure (0 = skip)")
    ap.add_argument("--simrank-scale", type=int, default=28, help="config-4 SimRank line on R-MAT of this scale (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    eng = gw.Engine(0)
    info = eng.device_info()
    print(f"[bench] {info}", file=sys.stderr, flush=True)

    shape = tuple(int(x) for x in args.shape.split(",")) if args.shape else None
    m = measure(eng, args.workload, args.steps, args.warmup, args.schedule, args.layout, compare=True, shape=shape)
    wl = m["wl"]
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(eng, wl, m["prog"], args.cpu_seconds)

    # random-gather ceiling: independent 4-byte gathers over 4 GiB
    gathers = 1 << 30
    gms = eng.gather_probe(4 << 30, gathers)
    rand_gbps = gathers * SECTOR / (gms * 1e-3) / 1e9
    # streaming copy of 1 GiB: sequential ceiling, and the known-size kernel
    # that calibrates FETCH_SIZE / WRITE_SIZE in profiles
    copy_bytes = 1 << 30
    cms = eng.copy_probe(copy_bytes)

    # HBM bytes per launch of the dominant kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # (profiles/), calibrated against probe kernels of known size
    traffic = traffic_copy = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{args.schedule}_{args.layout}.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        traffic, traffic_copy = t.get("bytes_per_launch"), t.get("bytes_per_launch_copy_calibrated")

    # post-processing on the headline graph: top-k of the per-vertex counts (the toplist caller)
    post = {}
    eng.run(gw.program(wl["kind"], wl["L"], wl["p"], m["n"], seed=wl["seed"], recorder=gw.REC_VERTEX_COUNT,
                       schedule=args.schedule, layout=args.layout))
    for k in (100, 1 << 20):
        eng.counts_topk(k)
        t = time.perf_counter()
        eng.counts_topk(k)
        post[f"counts_topk_{k}_ms"] = (time.perf_counter() - t) * 1e3
    print(f"[bench] counts top-k: {post}", file=sys.stderr, flush=True)

    # the same walks as a host walk pool (out-of-HBM walk pools), stepped in windows
    pool = {}
    if args.pool_window:
        for keep in (False, True):
            r = pool_measure(eng, wl, m["n"], m["prog"], args.pool_window, max(2, args.steps // 3), keep)
            pool["with_finals" if keep else "counts_only"] = r
            print(f"[bench] host walk pool (window {args.pool_window}, finals {keep}): "
                  f"{r['walk_steps_per_s'] / 1e9:.2f} G steps/s, {r['pcie_GBps']:.1f} GB/s over PCIe",
                  file=sys.stderr, flush=True)

    workloads = {}
    if args.secondary and args.secondary != args.workload:
        s = measure(eng, args.secondary, max(3, args.steps // 2), 1, args.schedule, args.layout, compare=False,
                    shape=shape)
        _, w_, st_, vi_, k_, _ = timed(eng, s["wl"], s["n"], "plain", "csr", 2, 1, source_method=gw.SRC_GLOBAL)
        _, wh_, sth_, _, _, _ = timed(eng, s["wl"], s["n"], "plain", "csr", 2, 1, source_method=gw.SRC_HASH)
        _, wg_, stg_, _, _, lg_ = timed(eng, s["wl"], s["n"], args.schedule, args.layout, 2, 1,
                                        source_method=gw.SRC_GLOBAL)
        last = s["last"]
        workloads[args.secondary] = {
            "desc": s["wl"]["desc"], "walk_steps_per_s": s["value"], "visits_per_s": s["visits"] * (s["value"] / s["steps"]),
            "walk_steps_per_pass": s["steps"], "visits_per_pass": s["visits"], "kernel_ms": s["kernel_ms"],
            "roofline_achieved_GBps": s["achieved"], "roofline_frac": s["achieved"] / PEAK_HBM_GBPS,
            "plain_csr_walk_steps_per_s": st_ / w_, "speedup_vs_plain": s["value"] / (st_ / w_),
            "plain_csr_hash_walk_steps_per_s": sth_ / wh_, "speedup_vs_plain_hash": s["value"] / (sth_ / wh_),
            "schedule": f"{args.schedule}-{args.layout}",
            "source_method": {gw.SRC_PARTITIONED: "partitioned", gw.SRC_GLOBAL: "global-sort",
                              gw.SRC_HASH: "hash"}.get(last["source_method"], "?"),
            "reduce_ms": last["reduce_ms"], "flushed_sources": last["flushed_sources"],
            "global_sort_walk_steps_per_s": stg_ / wg_, "global_sort_reduce_ms": lg_["reduce_ms"],
            "launch_shape": s["tuned"], "graph_replay": s.get("graph_replay"),
        }

    serving = {}
    if args.serve_qps:
        # continuous batching on the graph left resident by the last workload
        n_last = 1 << (WORKLOADS[args.secondary]["scale"] if workloads else wl["scale"])
        for q in args.serve_qps:
            r = serve_sim(eng, n_last, q, args.serve_seconds, hops=args.serve_hops, topk=args.serve_topk)
            serving[f"qps_{int(q)}"] = r
            print(f"[bench] serving {q:.0f} qps: p50 {r['p50_ms']:.2f} ms p99 {r['p99_ms']:.2f} ms, "
                  f"{r['walk_steps_per_s'] / 1e9:.2f} G steps/s", file=sys.stderr, flush=True)

    stream = {}
    if args.stream_cache_mib:
        wl_last = WORKLOADS[args.secondary] if workloads else wl
        stream = stream_measure(eng, wl_last["scale"], wl_last["seed"], args.stream_cache_mib << 20,
                                args.stream_queries, 2)

    simrank = {}
    if args.simrank_scale:
        simrank = simrank_measure(eng, args.simrank_scale, 4, 1024, 1000, 11, 0.6, 3)

    out = {
        "metric": "walk-steps/s",
        "value": m["value"],
        "unit": "walk-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["wall_s"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (R-MAT generated on the device)",
        "config": {"workload": args.workload, "desc": wl["desc"], "walk_steps_per_pass": m["steps"],
                   "schedule": args.schedule, "layout": args.layout, "device": info["name"]},
        "roofline": {"bound": "hbm", "achieved": m["achieved"], "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": m["achieved"] / PEAK_HBM_GBPS, "traffic": traffic,
                     "traffic_copy_calibrated": traffic_copy,
                     "kernel": f"step_{args.schedule}_kernel ({args.layout})", "kernel_ms": m["kernel_ms"],
                     "kernel_ms_median": m["kernel_ms_median"],
                     "algorithmic_bytes_per_launch": m["algorithmic_bytes"],
                     "bytes_per_step": m["algorithmic_bytes"] / max(1, m["steps"])},
        "random_access_ceiling": {"gathers_per_s": gathers / (gms * 1e-3), "GBps_at_64B": rand_gbps,
                                  "frac_of_ceiling": m["achieved"] / rand_gbps},
        "copy_GBps": 2 * copy_bytes / (cms * 1e-3) / 1e9,
        "variants": m["variants"],
        "launch_shape": m["tuned"],
        "graph_replay": m.get("graph_replay"),
        "workloads": workloads,
        "serving": serving,
        "out_of_core": stream,
        "workloads_simrank": simrank,
        "post_processing": post,
        "host_walk_pool": pool,
    }
    if "speedup_vs_plain" in m:
        out["speedup_vs_plain"] = m["speedup_vs_plain"]
    if cpu:
        out["cpu_baseline"] = cpu
    eng.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
