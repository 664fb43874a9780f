"""CPU oracle for the walk-update hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker (or as the timed CPU
baseline).  The shipped library (``graphwalker_amd``) never imports it.

``gw_oracle.c`` restates the reference's step programs
(src/walks/randomwalkwith{jump,stop,prob,restartwithjoint,restart}.hpp),
block loop structure (src/engine/graphwalker.hpp:257-311) and walk bucketing
(src/walks/walk.hpp:93-137) with Philox4x32-10 in place of glibc ``rand_r``.
``fp64.py`` holds the FP64 finite-horizon expectations the Monte-Carlo results
are compared with.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgw_oracle.so")

JUMP, STOP, PROB, RESTART_JOINT, RESTART = 0, 1, 2, 3, 4
MOVED, FINISHED = 0, 1

WALK_DTYPE = np.dtype([("id", "<u4"), ("source", "<u4"), ("vertex", "<u4"), ("hop", "<u4")])


class Program(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("L", C.c_uint32),
        ("stop_thr", C.c_uint64),
        ("nvertices", C.c_uint32),
        ("stream", C.c_uint32),
        ("seed", C.c_uint64),
    ]


def stop_threshold(p: float) -> int:
    """Coin threshold: a walk follows an edge iff r0 >= thr, r0 uniform u32.
    p = 0.15 reproduces ``rand_r/RAND_MAX > 0.15`` (randomwalkwithstop.hpp:35)."""
    if not 0.0 <= p <= 1.0:
        raise ValueError("probability out of range")
    return int(round(p * 2.0 ** 32))


def program(kind: int, L: int, p: float, nvertices: int, seed: int, stream: int = 0) -> Program:
    if L < 1 and kind != PROB:
        raise ValueError("L must be >= 1")
    return Program(kind, L, stop_threshold(p), nvertices, stream, seed)


def rmat_thresholds(a: float, b: float, c: float) -> tuple[int, int, int]:
    """Cumulative quadrant probabilities as 16-bit integer thresholds."""
    return (int(round(a * 65536)), int(round((a + b) * 65536)), int(round((a + b + c) * 65536)))


def _build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            _build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        u32, u64 = C.c_uint32, C.c_uint64
        L.gwo_philox4x32_10.argtypes = [P, P, P]
        L.gwo_philox_batch.argtypes = [P, u64, P, P]
        L.gwo_rmat_edges.argtypes = [u32, u64, u64, u64, u32, u32, u32, P, P]
        L.gwo_sources.argtypes = [u64, u32, u32, P]
        L.gwo_csr_build.argtypes = [u32, u64, P, P, P, P]
        L.gwo_csr_build.restype = C.c_int
        L.gwo_replay.argtypes = [P, P, P, u32, P, u64, P, P, P, P, P]
        L.gwo_trace.argtypes = [P, P, P, u32, P, P, P, u64, P]
        L.gwo_trace.restype = u64
        L.gwo_block_round.argtypes = [P, P, P, u32, P, u32, u32, P, u64, P, P, P]
        L.gwo_bucket_stable.argtypes = [P, u64, P, u32, P, P]
        L.gwo_fp64_stop_visits.argtypes = [u32, P, P, P, u32, P, C.c_double, u32, C.c_double, P]
        L.gwo_num_threads.restype = C.c_int
        L.gwo_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key) -> np.ndarray:
    """Philox4x32-10 of a (n, 4) uint32 counter array under a 2-word key."""
    ctr = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32).reshape(-1, 4))
    key = np.ascontiguousarray(np.asarray(key, dtype=np.uint32).reshape(2))
    out = np.empty_like(ctr)
    lib().gwo_philox_batch(_p(ctr), ctr.shape[0], _p(key), _p(out))
    return out


def rmat_edges(scale: int, m: int, seed: int, a=0.57, b=0.19, c=0.19, e_begin: int = 0):
    ta, tab, tabc = rmat_thresholds(a, b, c)
    src = np.empty(m, dtype=np.uint32)
    dst = np.empty(m, dtype=np.uint32)
    lib().gwo_rmat_edges(scale, e_begin, e_begin + m, seed, ta, tab, tabc, _p(src), _p(dst))
    return src, dst


def fp64_stop_visits(off: np.ndarray, toff: np.ndarray, tcol: np.ndarray, sources, walks: float, L: int,
                     p: float) -> np.ndarray:
    """FP64 expected visits of `walks` STOP walks (cap L) from each source:
    row k for sources[k] (at most 64), pulled over the transposed CSR (toff,
    tcol) with the forward offsets' out-degrees; fp64.expected_visits_sources
    for graphs too large for scipy."""
    src = np.ascontiguousarray(sources, dtype=np.uint32)
    K = src.shape[0]
    if not 1 <= K <= 64:
        raise ValueError("1..64 sources")
    off = np.ascontiguousarray(off, dtype=np.uint64)
    toff = np.ascontiguousarray(toff, dtype=np.uint64)
    tcol = np.ascontiguousarray(tcol, dtype=np.uint32)
    n = off.shape[0] - 1
    out = np.empty((n, K), dtype=np.float64)
    lib().gwo_fp64_stop_visits(n, _p(toff), _p(tcol), _p(off), K, _p(src), float(walks), L, float(p), _p(out))
    return out.T.copy()


def csr_build(n: int, src: np.ndarray, dst: np.ndarray):
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    off = np.empty(n + 1, dtype=np.uint64)
    col = np.empty(src.shape[0], dtype=np.uint32)
    rc = lib().gwo_csr_build(n, src.shape[0], _p(src), _p(dst), _p(off), _p(col))
    if rc != 0:
        raise ValueError(f"csr_build failed ({rc})")
    return off, col


def rmat_graph(scale: int, m: int, seed: int, a=0.57, b=0.19, c=0.19):
    src, dst = rmat_edges(scale, m, seed, a, b, c)
    return csr_build(1 << scale, src, dst)


def sources(seed: int, n: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    lib().gwo_sources(seed, n, count, _p(out))
    return out


def make_walks(vertices, sources_=None, hops=None, ids=None) -> np.ndarray:
    v = np.asarray(vertices, dtype=np.uint32)
    w = np.zeros(v.shape[0], dtype=WALK_DTYPE)
    w["vertex"] = v
    w["source"] = v if sources_ is None else np.asarray(sources_, dtype=np.uint32)
    w["hop"] = 0 if hops is None else np.asarray(hops, dtype=np.uint32)
    w["id"] = np.arange(v.shape[0], dtype=np.uint32) if ids is None else np.asarray(ids, dtype=np.uint32)
    return w


def replay(prog: Program, off, col, walks, counts: bool = True, per_walk: bool = False):
    """Run every walk to completion on the whole graph.  Returns a dict with
    ``counts`` (u32 per-vertex visits), ``steps``, ``visits`` and, with
    ``per_walk``, ``final`` / ``hash`` / ``walk_visits``."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    walks = np.ascontiguousarray(walks, dtype=WALK_DTYPE)
    n = off.shape[0] - 1
    nw = walks.shape[0]
    cnt = np.zeros(n, dtype=np.uint32) if counts else None
    fin = np.empty(nw, dtype=WALK_DTYPE) if per_walk else None
    hsh = np.empty(nw, dtype=np.uint64) if per_walk else None
    wv = np.empty(nw, dtype=np.uint32) if per_walk else None
    tot = np.zeros(2, dtype=np.uint64)
    lib().gwo_replay(C.byref(prog), _p(off), _p(col), n, _p(walks), nw, _p(cnt), _p(fin), _p(hsh), _p(wv), _p(tot))
    out = {"counts": cnt, "steps": int(tot[0]), "visits": int(tot[1])}
    if per_walk:
        out.update(final=fin, hash=hsh, walk_visits=wv)
    return out


def trace(prog: Program, off, col, walk, cap: int = 4096):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    w = np.ascontiguousarray(np.asarray([walk], dtype=WALK_DTYPE) if not isinstance(walk, np.ndarray) else walk.reshape(1))
    tv = np.empty(cap, dtype=np.uint32)
    th = np.empty(cap, dtype=np.uint32)
    fin = np.empty(1, dtype=WALK_DTYPE)
    n = lib().gwo_trace(C.byref(prog), _p(off), _p(col), off.shape[0] - 1, _p(w), _p(tv), _p(th), cap, _p(fin))
    k = min(n, cap)
    return tv[:k].copy(), th[:k].copy(), fin[0]


def block_round(prog: Program, off, col, block_first, b: int, walks, counts=None):
    """Advance bucket ``b`` in place; returns (status u8[], totals)."""
    bf = np.ascontiguousarray(block_first, dtype=np.uint32)
    status = np.empty(walks.shape[0], dtype=np.uint8)
    tot = np.zeros(2, dtype=np.uint64)
    lib().gwo_block_round(C.byref(prog), _p(off), _p(col), off.shape[0] - 1, _p(bf), bf.shape[0] - 1, b,
                          _p(walks), walks.shape[0], _p(status), _p(counts), _p(tot))
    return status, tot


def bucket_stable(walks, block_first):
    bf = np.ascontiguousarray(block_first, dtype=np.uint32)
    nb = bf.shape[0] - 1
    walks = np.ascontiguousarray(walks, dtype=WALK_DTYPE)
    counts = np.zeros(nb + 1, dtype=np.uint64)
    out = np.empty(walks.shape[0], dtype=WALK_DTYPE)
    lib().gwo_bucket_stable(_p(walks), walks.shape[0], _p(bf), nb, _p(out), _p(counts))
    return out[: int(counts[:nb].sum())], counts


def block_engine(prog: Program, off, col, block_first, walks, choose: str = "maxwalks"):
    """Restatement of graphwalker_engine::run (graphwalker.hpp:272-311): pick a
    block (max walks, walk.hpp:204-213), advance its bucket, re-bucket the
    walks that left, until no walk remains (hasFinishedWalk, randomwalk.hpp:87-91).
    Returns (counts, steps, visits, rounds)."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    bf = np.ascontiguousarray(block_first, dtype=np.uint32)
    nb = bf.shape[0] - 1
    n = off.shape[0] - 1
    counts = np.zeros(n, dtype=np.uint32)
    sorted_w, cnt = bucket_stable(walks, bf)
    buckets = []
    pos = 0
    for b in range(nb):
        k = int(cnt[b])
        buckets.append(sorted_w[pos:pos + k].copy())
        pos += k
    steps = visits = rounds = 0
    while sum(x.shape[0] for x in buckets) > 0:
        sizes = [x.shape[0] for x in buckets]
        b = int(np.argmax(sizes))
        cur = buckets[b]
        buckets[b] = np.empty(0, dtype=WALK_DTYPE)
        st, tot = block_round(prog, off, col, bf, b, cur, counts)
        steps += int(tot[0]); visits += int(tot[1]); rounds += 1
        moved = cur[st == MOVED]
        if moved.shape[0]:
            mw, mc = bucket_stable(moved, bf)
            p = 0
            for bb in range(nb):
                k = int(mc[bb])
                if k:
                    buckets[bb] = np.concatenate([buckets[bb], mw[p:p + k]])
                p += k
    return counts, steps, visits, rounds


def set_threads(t: int) -> None:
    lib().gwo_set_threads(t)


def num_threads() -> int:
    return lib().gwo_num_threads()


# --------------------------------------------------------------- reference apps
# The reference's own apps, compiled from its sources by build_ref.sh into
# _ref/ (a CPU baseline only; never the checker of a GPU result).

_REF_DIR = os.path.join(_HERE, "_ref")


def reference_app(name: str) -> str | None:
    """Path of a reference app built into oracle/_ref/, or None."""
    p = os.path.join(_REF_DIR, name)
    return p if os.access(p, os.X_OK) else None


def write_reference_graph(base: str, off: np.ndarray, col: np.ndarray) -> None:
    """A CSR in the reference's preprocessed layout, so its convert_if_notexists
    finds it (conversions.hpp:371-385, filename.hpp:16-51): the u64 absolute
    offsets <base>_GraphWalker/graphinfo/file_0.beg_pos (n + 1 entries), the u32
    columns file_0.csr, the one-file range filesize_1024GB.filerange and
    N.nvertices.  The reference computes its block range itself."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    n = off.shape[0] - 1
    root = base + "_GraphWalker"
    os.makedirs(os.path.join(root, "graphinfo"), exist_ok=True)
    off.tofile(os.path.join(root, "graphinfo", "file_0.beg_pos"))
    col[: int(off[-1])].tofile(os.path.join(root, "graphinfo", "file_0.csr"))
    with open(os.path.join(root, "filesize_1024GB.filerange"), "w") as f:
        f.write(f"0\n{n}\n")
    with open(os.path.join(root, "N.nvertices"), "w") as f:
        f.write(f"{n}\n")


def _run_reference(app_name: str, base: str, options: dict, timeout: float) -> str:
    """Runs an app of oracle/_ref/ on a graph written by write_reference_graph,
    in the graph's directory, with the given command-line options; returns
    its stdout."""
    app = reference_app(app_name)
    if app is None:
        raise FileNotFoundError(f"oracle/_ref/{app_name} is not built")
    cwd = os.path.dirname(base)
    # the reference's option parser insists on a config file (configfile.hpp:81-90);
    # every option here comes from the command line
    os.makedirs(os.path.join(cwd, "conf"), exist_ok=True)
    with open(os.path.join(cwd, "conf", "graphwalker.local.cnf"), "w") as f:
        f.write("# options are given on the command line\n")
    args = [app, "file", base]
    for k, v in options.items():
        args += [k, str(v)]
    r = subprocess.run(args, cwd=cwd, check=True, timeout=timeout, stdout=subprocess.PIPE,
                       stderr=subprocess.DEVNULL, text=True)
    return r.stdout


def run_reference_rawrandomwalks(base: str, n: int, R: int, L: int, threads: int, blocksize_kb: int,
                                 nmblocks: int, timeout: float) -> dict:
    """Runs the reference rawrandomwalks app (R walks from uniform sources,
    L jump steps each) and returns its own metrics (metrics.hpp timers):
    00_runtime (the run after walk placement) and 5_exec_updates (the step
    phase), in seconds."""
    _run_reference("rawrandomwalks", base, {"N": n, "R": R, "L": L, "execthreads": threads,
                                            "blocksize_kb": blocksize_kb, "nmblocks": nmblocks, "prob": 0.2},
                   timeout)
    out = {}
    with open(os.path.join(os.path.dirname(base), "graphwalker_metrics.txt")) as f:
        for line in f:
            key, _, val = line.strip().partition("=")
            for k in ("00_runtime", "5_exec_updates", "0_startWalks"):
                if key.endswith("." + k):
                    out[k] = float(val)
    return out


def run_reference_visit_counts(base: str, n: int, source: int, walks: int, L: int, threads: int,
                               timeout: float) -> np.ndarray:
    """The reference engine's single-source PPR visit vector: `walks` walks of
    its RandomWalkwithStop program (stop 0.15, cap L) from `source`, counted
    per vertex by oracle/ref_visit_counts.cpp (one block, all in memory)."""
    out = base + ".visits"
    _run_reference("visit_counts", base, {"N": n, "source": source, "walks": walks, "L": L, "out": out,
                                          "execthreads": threads, "blocksize_kb": 1 << 20, "nmblocks": 1},
                   timeout)
    return np.fromfile(out, dtype=np.uint32)


def run_reference_simrank(base: str, a: int, b: int, R: int, L: int, threads: int, timeout: float) -> float:
    """The reference simrank app's score for (a, b): R walks of its
    RandomWalkwithRestartwithJoint program from each vertex, first meetings
    weighed by 0.8^offset (simrank.cpp:71-84), one block all in memory."""
    text = _run_reference("simrank", base, {"a": a, "b": b, "R": R, "L": L, "execthreads": threads,
                                            "blocksize_kb": 1 << 20, "nmblocks": 1, "prob": 0.2}, timeout)
    for line in text.splitlines():
        if line.startswith(f"SimRank for {a} and {b} = "):
            return float(line.rsplit("=", 1)[1])
    raise RuntimeError("the reference simrank app printed no score")


def run_reference_graphlet(base: str, n: int, R: int, L: int, threads: int, timeout: float) -> int:
    """The reference graphlet app's closed-walk count (cnt_ok): R stop walks
    of cap L from rand() % N sources, counting visits to the source at hop
    L-1 (graphlet.cpp:46-52), one block all in memory."""
    text = _run_reference("graphlet", base, {"N": n, "R": R, "L": L, "execthreads": threads,
                                             "blocksize_kb": 1 << 20, "nmblocks": 1, "prob": 0.2}, timeout)
    for line in text.splitlines():
        if line.startswith("cnt_ok:"):
            return int(line.split()[-1])
    raise RuntimeError("the reference graphlet app printed no count")


def run_reference_msppr_sources(base: str, sources: np.ndarray, wps: int, L: int, threads: int, timeout: float,
                                blocksize_kb: int, nmblocks: int = 1) -> dict:
    """The reference engine on batched multi-source PPR: walks_per_source
    stop walks (0.15, cap L) from each vertex of `sources`, recorded as msppr
    records them (oracle/ref_msppr_sources.cpp).  Returns its own timers
    (00_runtime, 5_exec_updates, 0_startWalks, in seconds), the visit total,
    and source 0's (vertex, count) distribution.  blocksize_kb: the engine's
    block size (a block at least the whole column file keeps the graph in
    one in-memory block)."""
    src_path = base + ".sources"
    np.ascontiguousarray(sources, dtype=np.uint32).tofile(src_path)
    out = base + ".top"
    text = _run_reference("msppr_sources", base, {"sources": src_path, "walkspersource": wps, "L": L, "out": out,
                                                  "execthreads": threads, "blocksize_kb": blocksize_kb,
                                                  "nmblocks": nmblocks, "prob": 0.2}, timeout)
    res = {}
    with open(os.path.join(os.path.dirname(base), "graphwalker_metrics.txt")) as f:
        for line in f:
            key, _, val = line.strip().partition("=")
            for k in ("00_runtime", "5_exec_updates", "0_startWalks"):
                if key.endswith("." + k):
                    res[k] = float(val)
    for line in text.splitlines():
        if line.startswith("visits:"):
            res["visits"] = int(line.split()[-1])
    pairs = np.fromfile(out, dtype=np.uint32).reshape(-1, 2)
    res["source0_vertex"], res["source0_count"] = pairs[:, 0].copy(), pairs[:, 1].copy()
    return res
