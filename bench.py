"""Walk-steps/s of the walk-update hot path on one MI355X.

Default workload: raw random walks on R-MAT scale 24 (16,777,216 vertices,
edge factor 16 = 268,435,456 edges, a/b/c/d = .57/.19/.19/.05, seed 2), one
walk per vertex, fixed length 10, no restart (dangling vertices teleport):
167,772,160 walk-steps per pass.  A bench "step" is one pass of the engine over
all walks (gw_run), graph and walk records resident in HBM before timing.

Prints one JSON line (rank 0).  The CPU baseline is the oracle's OpenMP
replay of the same walk programs on a bounded sample of the same walks.
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
    # name: (scale, edge_factor, seed, kind, L, p, recorder, block_vertices)
    "rawwalks-rmat24": dict(scale=24, ef=16, seed=2, kind=gw.JUMP, L=10, p=0.0, recorder=gw.REC_NONE,
                            block_vertices=1 << 20,
                            desc="R-MAT scale 24 ef 16 seed 2, 1 walk/vertex, L=10, jump p=0"),
    "rawwalks-rmat20": dict(scale=20, ef=16, seed=2, kind=gw.JUMP, L=10, p=0.0, recorder=gw.REC_NONE,
                            block_vertices=1 << 16,
                            desc="R-MAT scale 20 ef 16 seed 2, 1 walk/vertex, L=10, jump p=0 (reduced)"),
}


def bytes_per_step(recorder: int, layout: int) -> int:
    """Algorithmic random-access bytes per walk-step, at 64-B sector
    granularity: the CSR layout reads one sector for the offset pair and one
    for the column; the successor-header layout reads one 16-B edge record
    (one sector).  A per-visit count adds one sector (its atomic)."""
    reads = 2 if layout == gw.LAYOUT_CSR else 1
    return SECTOR * (reads + (1 if recorder & gw.REC_VERTEX_COUNT else 0))


def cpu_baseline(eng: gw.Engine, wl: dict, prog: gw.Program, target_s: float) -> dict:
    import oracle as O

    off, col = eng.graph_read()
    op = O.Program(prog.kind, prog.L, prog.stop_thr, prog.nvertices, prog.stream, prog.seed)
    n = off.shape[0] - 1
    threads = O.num_threads()
    nw = 1 << 18
    elapsed = 0.0
    steps = 0
    while True:
        walks = O.make_walks(np.arange(nw, dtype=np.uint32) % n)
        t0 = time.perf_counter()
        r = O.replay(op, off, col, walks, counts=bool(prog.recorder & gw.REC_VERTEX_COUNT))
        elapsed = time.perf_counter() - t0
        steps = r["steps"]
        if elapsed >= target_s / 4 or nw >= n:
            break
        nw = min(n, int(nw * max(2.0, target_s / max(elapsed, 1e-3))))
    return {"value": steps / elapsed, "unit": "walk-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle OpenMP replay of walks 0..{nw - 1} (one per vertex) of the same program on the "
                      f"same device-generated graph, {steps} steps in {elapsed:.2f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rawwalks-rmat24", choices=sorted(WORKLOADS))
    ap.add_argument("--schedule", default="persistent", choices=["plain", "persistent", "blocked"])
    ap.add_argument("--layout", default="ext", choices=["csr", "ext"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    wl = WORKLOADS[args.workload]
    eng = gw.Engine(0)
    info = eng.device_info()
    print(f"[bench] {info}", file=sys.stderr, flush=True)
    n = 1 << wl["scale"]
    m = n * wl["ef"]
    t0 = time.perf_counter()
    eng.graph_rmat(wl["scale"], m, wl["seed"])
    eng.walks_init_per_vertex(1)
    print(f"[bench] graph + walks ready in {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)

    bf = np.arange(0, n + 1, wl["block_vertices"], dtype=np.uint32)
    eng.graph_blocks(bf)

    def make(sched, layout):
        return gw.program(wl["kind"], wl["L"], wl["p"], n, seed=wl["seed"], recorder=wl["recorder"], schedule=sched,
                          layout=layout)

    def timed(sched, layout, K, W):
        prog = make(sched, layout)
        for _ in range(W):
            eng.run(prog)
        t = time.perf_counter()
        kms, steps, st = [], 0, None
        for _ in range(K):
            st = eng.run(prog)
            kms.append(st["step_ms"])
            steps += st["steps"]
        wall = time.perf_counter() - t
        return prog, wall, steps, kms, st

    # the other schedules/layouts of the same workload, for comparison
    variants = {}
    for sched, layout in (("plain", "csr"), ("persistent", "csr"), ("persistent", "ext"), ("blocked", "ext")):
        if (sched, layout) == (args.schedule, args.layout):
            continue
        _, w_, s_, k_, st_ = timed(sched, layout, max(2, args.steps // 3), 1)
        variants[f"{sched}-{layout}"] = {"walk_steps_per_s": s_ / w_, "kernel_ms": float(np.mean(k_)),
                                         "rounds": st_["rounds"], "bucket_ms": st_["bucket_ms"]}
        print(f"[bench] {sched}-{layout}: {s_ / w_ / 1e9:.2f} G steps/s", file=sys.stderr, flush=True)
    prog, wall, steps, kms, st = timed(args.schedule, args.layout, args.steps, args.warmup)
    steps_per_pass = steps // args.steps
    value = steps / wall
    kavg = float(np.mean(kms))
    bstep = bytes_per_step(wl["recorder"], st["layout"])
    achieved = steps_per_pass * bstep / (kavg * 1e-3) / 1e9
    print(f"[bench] {args.schedule}-{args.layout}: {value / 1e9:.2f} G steps/s", file=sys.stderr, flush=True)

    # random-gather ceiling: independent 4-byte gathers over 4 GiB
    gathers = 1 << 30
    gms = eng.gather_probe(4 << 30, gathers)
    rand_gbps = gathers * SECTOR / (gms * 1e-3) / 1e9
    # streaming copy of 1 GiB: sequential ceiling, and the known-size kernel
    # that calibrates FETCH_SIZE / WRITE_SIZE in profiles
    copy_bytes = 1 << 30
    cms = eng.copy_probe(copy_bytes)

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{args.schedule}_{args.layout}.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")

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
        "config": {"workload": args.workload, "desc": wl["desc"], "walks": n, "walk_steps_per_pass": steps_per_pass,
                   "schedule": args.schedule, "layout": args.layout, "device": info["name"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBPS, "traffic": traffic,
                     "kernel": f"step_{args.schedule}_kernel ({args.layout})", "kernel_ms": kavg,
                     "bytes_per_step": bstep},
        "random_access_ceiling": {"gathers_per_s": gathers / (gms * 1e-3), "GBps_at_64B": rand_gbps,
                                  "frac_of_ceiling": achieved / rand_gbps},
        "copy_GBps": 2 * copy_bytes / (cms * 1e-3) / 1e9,
        "variants": variants,
    }
    if "plain-csr" in variants:
        out["speedup_vs_plain"] = value / variants["plain-csr"]["walk_steps_per_s"]
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(eng, wl, prog, args.cpu_seconds)
    eng.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
