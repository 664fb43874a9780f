"""graphwalker_amd -- MI355X-native walk-update engine (Python host binding).

Thin ctypes layer over ``libgwhip.so`` (include/gwhip.h).  Every call goes to
the HIP library; there is no CPU fallback: if the library is missing or no GPU
is present the constructor raises.

The names follow the GraphWalker engine they replace: a walk *program*
(RandomWalkwithJump / Stop / Prob / RestartwithJoint / Restart,
src/walks/randomwalk*.hpp), *blocks* of vertices (graphwalker.hpp:146-166),
*walk records* (walk.hpp:69-91) and the engine *run* (graphwalker.hpp:272-311).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgwhip.so")

# walk programs (include/gwhip.h)
JUMP, STOP, PROB, RESTART_JOINT, RESTART = 0, 1, 2, 3, 4
KIND_NAMES = {"jump": JUMP, "stop": STOP, "prob": PROB, "restart_joint": RESTART_JOINT, "restart": RESTART}
# recorders
REC_NONE, REC_VERTEX_COUNT, REC_WALK_SUMMARY, REC_EVENTS, REC_SOURCE_COUNT = 0, 1, 2, 4, 8
REC_NO_FINAL = 16  # skip the final walk records (flag, with any recorder)
REC_EVENTS_LAST = 32  # with REC_EVENTS: visits past event_cap overwrite the log's last slot
# schedules
SCHED_AUTO, SCHED_PLAIN, SCHED_PERSISTENT, SCHED_BLOCKED = 0, 1, 2, 3
SCHED_NAMES = {"auto": SCHED_AUTO, "plain": SCHED_PLAIN, "persistent": SCHED_PERSISTENT, "blocked": SCHED_BLOCKED}
# graph layouts the step kernels read
LAYOUT_AUTO, LAYOUT_CSR, LAYOUT_EXT = 0, 1, 2
LAYOUT_NAMES = {"auto": LAYOUT_AUTO, "csr": LAYOUT_CSR, "ext": LAYOUT_EXT}
# per-source count aggregation
SRC_SORT, SRC_HASH, SRC_GLOBAL, SRC_PARTITIONED = 0, 1, 2, 3
POOL_KEEP_FINAL = 1
EDGES_DROP_SELF_LOOPS, EDGES_SORT_ROWS = 1, 2
BLOCKS_MAX_WALKS, BLOCKS_MIN_STEP, BLOCKS_MAX_WEIGHT = 0, 1, 2

WALK_DTYPE = np.dtype([("id", "<u4"), ("source", "<u4"), ("vertex", "<u4"), ("hop", "<u4")])


class GwError(RuntimeError):
    pass


class _Cfg(C.Structure):
    _fields_ = [("device", C.c_int), ("reserved", C.c_uint32), ("hbm_budget_bytes", C.c_uint64)]


class Program(C.Structure):
    """gw_program: one walk program and how to run it."""

    _fields_ = [
        ("kind", C.c_uint32),
        ("L", C.c_uint32),
        ("stop_thr", C.c_uint64),
        ("nvertices", C.c_uint32),
        ("stream", C.c_uint32),
        ("seed", C.c_uint64),
        ("recorder", C.c_uint32),
        ("schedule", C.c_uint32),
        ("block_threads", C.c_uint32),
        ("waves_per_simd", C.c_uint32),
        ("layout", C.c_uint32),
        ("event_cap", C.c_uint32),
        ("table_log2", C.c_uint32),
        ("source_method", C.c_uint32),
        ("walk_chunk", C.c_uint32),
    ]


# gw_query as a numpy record: (qid, source, walks)
QUERY_DTYPE = np.dtype([("qid", np.uint64), ("source", np.uint32), ("walks", np.uint32)])


class Query(C.Structure):
    _fields_ = [("qid", C.c_uint64), ("source", C.c_uint32), ("walks", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64),
        ("visits", C.c_uint64),
        ("live_walks", C.c_uint64),
        ("admitted", C.c_uint32),
        ("completed", C.c_uint32),
        ("active_queries", C.c_uint32),
        ("queued", C.c_uint32),
        ("step_ms", C.c_double),
        ("h2d_bytes", C.c_uint64),
        ("pages_loaded", C.c_uint32),
        ("pages_resident", C.c_uint32),
        ("sampled_reads", C.c_uint64),
        ("sampled_resident", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64),
        ("visits", C.c_uint64),
        ("rounds", C.c_uint64),
        ("launches", C.c_uint64),
        ("step_ms", C.c_double),
        ("total_ms", C.c_double),
        ("bucket_ms", C.c_double),
        ("schedule", C.c_uint32),
        ("layout", C.c_uint32),
        ("reduce_ms", C.c_double),
        ("source_method", C.c_uint32),
        ("flushed_sources", C.c_uint32),
        ("h2d_bytes", C.c_uint64),
        ("blocks_loaded", C.c_uint32),
        ("cache_slots", C.c_uint32),
        ("copy_ms", C.c_double),
        ("pool_windows", C.c_uint64),
        ("pool_bytes", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def stop_threshold(p: float) -> int:
    """A walk follows an edge iff its 32-bit coin r0 >= p * 2^32; p = 0.15 is
    the reference's ``rand_r/RAND_MAX > 0.15`` (randomwalkwithstop.hpp:35)."""
    if not 0.0 <= p <= 1.0:
        raise ValueError("probability out of range")
    return int(round(p * 2.0 ** 32))


def program(kind, L: int, p: float = 0.15, nvertices: int = 0, seed: int = 0, recorder: int = REC_NONE,
            schedule=SCHED_AUTO, stream: int = 0, block_threads: int = 0, waves_per_simd: int = 0,
            layout=LAYOUT_AUTO, event_cap: int = 0, table_log2: int = 0, source_method: int = SRC_SORT,
            walk_chunk: int = 0) -> Program:
    if isinstance(kind, str):
        kind = KIND_NAMES[kind]
    if isinstance(schedule, str):
        schedule = SCHED_NAMES[schedule]
    if isinstance(layout, str):
        layout = LAYOUT_NAMES[layout]
    return Program(kind, L, stop_threshold(p), nvertices, stream, seed, recorder, schedule, block_threads,
                   waves_per_simd, layout, event_cap, table_log2, source_method, walk_chunk)


def rmat_thresholds(a: float, b: float, c: float) -> tuple[int, int, int]:
    """Cumulative R-MAT quadrant probabilities as 16-bit thresholds."""
    return (int(round(a * 65536)), int(round((a + b) * 65536)), int(round((a + b + c) * 65536)))


_lib = None


def block_cut(off: np.ndarray, blocksize_kb: int) -> np.ndarray:
    """Block table of a CSR by the reference's cut rule (compute_block,
    preprocess/conversions.hpp:300-364): a block closes before the vertex
    whose row would take it past blocksize_kb KB of columns, and a row larger
    than a block gets a block of its own."""
    off = np.asarray(off, dtype=np.uint64)
    n = off.shape[0] - 1
    mneb = blocksize_kb * 1024 // 4
    blocks = [0]
    start, v = 0, 1
    while v <= n:
        # first vertex v' >= v whose row end passes start + mneb
        w = int(np.searchsorted(off, np.uint64(start + mneb), side="right"))
        if w < v:
            w = v
        if w > n:
            break
        if int(off[w]) - int(off[w - 1]) > mneb:
            blocks += [w - 1, w]
            start = int(off[w])
        else:
            blocks.append(w - 1)
            start = int(off[w - 1])
        v = w + 1
    blocks.append(n)
    return np.asarray(blocks, dtype=np.uint32)


def write_graph_files(base: str, off: np.ndarray, col: np.ndarray, blocksize_kb: int) -> int:
    """A CSR in the preprocessed on-disk layout the GraphWalker engine reads
    (filename.hpp:16-51, conversions.hpp:181-298): <base>_GraphWalker/
    graphinfo/file_0.beg_pos (u64 absolute offsets) and file_0.csr (u32
    columns), N.nvertices, filesize_1024GB.filerange and the block table
    blocksize_<KB>KB.blockrange.  Returns the number of blocks."""
    d = base + "_GraphWalker"
    os.makedirs(os.path.join(d, "graphinfo"), exist_ok=True)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    np.ascontiguousarray(col, dtype=np.uint32).tofile(os.path.join(d, "graphinfo", "file_0.csr"))
    off.tofile(os.path.join(d, "graphinfo", "file_0.beg_pos"))
    n = off.shape[0] - 1
    with open(os.path.join(d, "filesize_1024GB.filerange"), "w") as f:
        f.write(f"0\n{n}\n")
    with open(os.path.join(d, "N.nvertices"), "w") as f:
        f.write(f"{n}\n")
    blocks = block_cut(off, blocksize_kb)
    with open(os.path.join(d, f"blocksize_{blocksize_kb}KB.blockrange"), "w") as f:
        f.write("".join(f"{int(b)}\n" for b in blocks))
    return blocks.shape[0] - 1


def sources(seed: int, n: int, count: int) -> np.ndarray:
    """Uniform source vertices reproducible from (seed, index) (gw_sources)."""
    out = np.empty(count, dtype=np.uint32)
    if lib().gw_sources(seed, n, count, _ptr(out)) != 0:
        raise GwError("gw_sources: bad arguments")
    return out


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GwError(f"{LIB_PATH} is not built; run `python -m graphwalker_amd.build`")
    L = C.CDLL(LIB_PATH)
    P, u32, u64, i = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    sig = {
        "gw_create": ([P, P], i),
        "gw_destroy": ([P], None),
        "gw_last_error": ([P], C.c_char_p),
        "gw_device_info": ([P, C.c_char_p, C.c_size_t, P, P], i),
        "gw_graph_set": ([P, P, P, u32, u64], i),
        "gw_graph_set_host": ([P, P, P, u32, u64, u64], i),
        "gw_graph_transpose": ([P], i),
        "gw_walks_init_pairs": ([P, P, P, u32, u32, u32], i),
        "gw_meeting_counts": ([P, u32, u32, u32, P], i),
        "gw_host_alloc": ([P, u64, P], i),
        "gw_host_free": ([P, P], i),
        "gw_graph_rmat": ([P, u32, u64, u64, u32, u32, u32], i),
        "gw_graph_set_edges": ([P, P, P, u64, u32, u32], i),
        "gw_graph_rmat_host": ([P, u32, u64, u64, u32, u32, u32, u64, u64, P], i),
        "gw_graph_blocks": ([P, P, u32], i),
        "gw_graph_info": ([P, P, P, P], i),
        "gw_graph_read": ([P, P, P], i),
        "gw_walks_load": ([P, P, u64], i),
        "gw_walks_load_host": ([P, P, u64, u64, u32], i),
        "gw_stream_policy": ([P, u32, C.c_double, u64], i),
        "gw_walks_init_per_vertex": ([P, u32], i),
        "gw_walks_init_sources": ([P, P, P, u32, u32], i),
        "gw_walks_read": ([P, P, u64], i),
        "gw_walks_current": ([P, P, u64], i),
        "gw_run": ([P, P, P], i),
        "gw_run_capture": ([P, P], i),
        "gw_run_replay": ([P, P], i),
        "gw_run_noalloc": ([P, P, P, u64, P, P, P], i),
        "gw_autotune": ([P, P, u32, P, P], i),
        "gw_bucket": ([P, P], i),
        "gw_counts_read": ([P, P, u32], i),
        "gw_summary_read": ([P, P, P, u64], i),
        "gw_events_read": ([P, P, u64, u32], i),
        "gw_source_counts_read": ([P, P, P, P, u64, P], i),
        "gw_source_topk": ([P, u32, P, P, P, u64, P], i),
        "gw_counts_topk": ([P, u64, P, P, P], i),
        "gw_gather_probe": ([P, u64, u64, P], i),
        "gw_copy_probe": ([P, u64, P], i),
        "gw_h2d_probe": ([P, u64, P], i),
        "gw_host_gather_probe": ([P, u64, u64, P], i),
        "gw_sources": ([u64, u32, u32, P], i),
        "gw_serve_open": ([P, P, u64, u64, u32, u32], i),
        "gw_serve_submit": ([P, P, u32], i),
        "gw_serve_step": ([P, P], i),
        "gw_serve_poll": ([P, P, P], i),
        "gw_serve_result": ([P, u64, P, P, u64], i),
        "gw_serve_results": ([P, u32, P, P, P, P, u64, P], i),
        "gw_serve_topk": ([P, u32], i),
        "gw_serve_close": ([P], i),
        "gw_serve_cache": ([P, u32, u32], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Engine:
    """One engine context on one GPU (gw_ctx)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._L = lib()
        self._ctx = C.c_void_p()
        cfg = _Cfg(device, 0, hbm_budget_bytes)
        rc = self._L.gw_create(C.byref(self._ctx), C.byref(cfg))
        if rc != 0:
            raise GwError(f"gw_create failed ({rc}): no usable HIP device {device}")
        self.n = 0
        self.m = 0
        self.nwalks = 0
        self._host_graph = None
        self._meet_buf = None  # pinned output of meeting_counts, grown as needed

    def close(self) -> None:
        if self._ctx:
            self._L.gw_destroy(self._ctx)  # frees gw_host_alloc memory too
            self._ctx = C.c_void_p()
            self._host_graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self._L.gw_last_error(self._ctx)
            raise GwError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ---------------------------------------------------------------- device
    def device_info(self) -> dict:
        name = C.create_string_buffer(256)
        cu = C.c_uint32()
        mem = C.c_uint64()
        self._check(self._L.gw_device_info(self._ctx, name, 256, C.byref(cu), C.byref(mem)), "gw_device_info")
        return {"name": name.value.decode(), "cu_count": cu.value, "hbm_bytes": mem.value}

    # ----------------------------------------------------------------- graph
    def graph_set(self, off: np.ndarray, col: np.ndarray) -> None:
        off = np.ascontiguousarray(off, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        n = off.shape[0] - 1
        self._check(self._L.gw_graph_set(self._ctx, _ptr(off), _ptr(col), n, col.shape[0]), "gw_graph_set")
        self.n, self.m = n, col.shape[0]

    def graph_set_edges(self, src: np.ndarray, dst: np.ndarray, n: int = 0, drop_self_loops: bool = False,
                        sort_rows: bool = False) -> None:
        """CSR built on the device from an edge list (rows keep the list's
        order unless sort_rows); n = 0: the largest vertex id + 1."""
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        if src.shape != dst.shape:
            raise GwError("graph_set_edges: src and dst differ in length")
        flags = (EDGES_DROP_SELF_LOOPS if drop_self_loops else 0) | (EDGES_SORT_ROWS if sort_rows else 0)
        self._check(self._L.gw_graph_set_edges(self._ctx, _ptr(src), _ptr(dst), src.shape[0], n, flags),
                    "gw_graph_set_edges")
        info = self.graph_info()
        self.n, self.m = info["n"], info["m"]

    def host_empty(self, shape, dtype) -> np.ndarray:
        """A numpy array in pinned host memory (gw_host_alloc), freed with the
        engine.  A host-resident graph built in such arrays is used in place."""
        dtype = np.dtype(dtype)
        count = int(np.prod(shape))
        ptr = C.c_void_p()
        self._check(self._L.gw_host_alloc(self._ctx, max(1, count * dtype.itemsize), C.byref(ptr)), "gw_host_alloc")
        buf = (C.c_char * max(1, count * dtype.itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)

    def graph_set_host(self, off: np.ndarray, col: np.ndarray, cache_bytes: int) -> None:
        """Out-of-core graph: the CSR stays in host memory and runs step it
        through a device block cache of at most cache_bytes."""
        off = np.ascontiguousarray(off, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        n = off.shape[0] - 1
        self._check(self._L.gw_graph_set_host(self._ctx, _ptr(off), _ptr(col), n, col.shape[0], cache_bytes),
                    "gw_graph_set_host")
        self.n, self.m = n, col.shape[0]
        self._host_graph = (off, col)  # the engine may use them in place

    def graph_rmat(self, scale: int, m: int, seed: int, a: float = 0.57, b: float = 0.19, c: float = 0.19) -> None:
        ta, tab, tabc = rmat_thresholds(a, b, c)
        self._check(self._L.gw_graph_rmat(self._ctx, scale, m, seed, ta, tab, tabc), "gw_graph_rmat")
        self.n, self.m = 1 << scale, m

    def graph_rmat_host(self, scale: int, m: int, seed: int, cache_bytes: int, work_bytes: int = 0,
                        a: float = 0.57, b: float = 0.19, c: float = 0.19) -> float:
        """The graph_rmat graph built into pinned host memory (host-resident,
        device block cache of cache_bytes); returns the build time in ms."""
        ta, tab, tabc = rmat_thresholds(a, b, c)
        ms = C.c_double()
        self._check(self._L.gw_graph_rmat_host(self._ctx, scale, m, seed, ta, tab, tabc, cache_bytes, work_bytes,
                                                C.byref(ms)), "gw_graph_rmat_host")
        self.n, self.m = 1 << scale, m
        self._host_graph = None
        return ms.value

    def graph_transpose(self) -> None:
        """Reverse every edge of the device-resident graph."""
        self._check(self._L.gw_graph_transpose(self._ctx), "gw_graph_transpose")

    def graph_info(self) -> dict:
        n, m, nb = C.c_uint32(), C.c_uint64(), C.c_uint32()
        self._check(self._L.gw_graph_info(self._ctx, C.byref(n), C.byref(m), C.byref(nb)), "gw_graph_info")
        return {"n": n.value, "m": m.value, "nblocks": nb.value}

    def graph_blocks(self, block_first) -> None:
        bf = np.ascontiguousarray(block_first, dtype=np.uint32)
        self._check(self._L.gw_graph_blocks(self._ctx, _ptr(bf), bf.shape[0] - 1), "gw_graph_blocks")

    def graph_read(self, off: np.ndarray | None = None, col: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
        """The resident graph's CSR, into new arrays or the given ones (e.g.
        host_empty arrays, to move a device-built graph to host memory)."""
        off = np.empty(self.n + 1, dtype=np.uint64) if off is None else off
        col = np.empty(self.m, dtype=np.uint32) if col is None else col
        if off.dtype != np.uint64 or off.shape != (self.n + 1,) or col.dtype != np.uint32 or col.shape != (self.m,):
            raise GwError("graph_read: output arrays must be uint64[n + 1] and uint32[m]")
        if not (off.flags.c_contiguous and col.flags.c_contiguous):
            raise GwError("graph_read: output arrays must be contiguous")
        self._check(self._L.gw_graph_read(self._ctx, _ptr(off), _ptr(col)), "gw_graph_read")
        return off, col

    # ----------------------------------------------------------------- walks
    def walks_load(self, walks: np.ndarray) -> None:
        walks = np.ascontiguousarray(walks, dtype=WALK_DTYPE)
        self._check(self._L.gw_walks_load(self._ctx, _ptr(walks), walks.shape[0]), "gw_walks_load")
        self.nwalks = walks.shape[0]

    def stream_policy(self, policy: int, prob: float = 0.2, seed: int = 0) -> None:
        """Block choice of out-of-core load passes: BLOCKS_MAX_WALKS (default),
        BLOCKS_MIN_STEP (the reference's chooseBlock(prob)) or BLOCKS_MAX_WEIGHT."""
        self._check(self._L.gw_stream_policy(self._ctx, policy, prob, seed), "gw_stream_policy")

    def walks_load_host(self, walks: np.ndarray, window: int, keep_final: bool = True) -> None:
        """Host walk pool: the records (walks[i].id == i) stay in host memory
        -- an array from host_empty is copied from directly -- and runs step
        them `window` walks at a time.  The array is kept alive here."""
        if walks.dtype != WALK_DTYPE or not walks.flags.c_contiguous:
            walks = np.ascontiguousarray(walks, dtype=WALK_DTYPE)
        self._check(self._L.gw_walks_load_host(self._ctx, _ptr(walks), walks.shape[0], window,
                                               POOL_KEEP_FINAL if keep_final else 0), "gw_walks_load_host")
        self._pool = walks
        self.nwalks = walks.shape[0]

    def walks_init_per_vertex(self, per_vertex: int = 1) -> None:
        self._check(self._L.gw_walks_init_per_vertex(self._ctx, per_vertex), "gw_walks_init_per_vertex")
        self.nwalks = self.n * per_vertex

    def walks_init_sources(self, sources, walks_per_source: int, source_field=None) -> None:
        s = np.ascontiguousarray(sources, dtype=np.uint32)
        f = None if source_field is None else np.ascontiguousarray(source_field, dtype=np.uint32)
        self._check(self._L.gw_walks_init_sources(self._ctx, _ptr(s), _ptr(f), s.shape[0], walks_per_source),
                    "gw_walks_init_sources")
        self.nwalks = s.shape[0] * walks_per_source

    def walks_init_pairs(self, a, b, R: int, L: int) -> None:
        """SimRank placement: R walks from a[k] then R from b[k] per pair, walk
        i at hop i*L (ids 2kR + i and 2kR + R + i)."""
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        self._check(self._L.gw_walks_init_pairs(self._ctx, _ptr(a), _ptr(b), a.shape[0], R, L), "gw_walks_init_pairs")
        self.nwalks = 2 * a.shape[0] * R

    def meeting_counts(self, npairs: int, R: int, L: int) -> np.ndarray:
        """First-meeting counts per pair and offset (npairs x L) of the last
        GW_RESTART_JOINT run with GW_REC_EVENTS over walks_init_pairs walks."""
        n = npairs * L
        if self._meet_buf is None or self._meet_buf.shape[0] < n:
            self._meet_buf = self.host_empty((n,), np.uint32)  # pinned: the copy runs at link speed
        buf = self._meet_buf[:n]
        self._check(self._L.gw_meeting_counts(self._ctx, npairs, R, L, _ptr(buf)), "gw_meeting_counts")
        return buf.reshape(npairs, L).copy()

    def walks_final(self) -> np.ndarray:
        """Final walk records, indexed by walk id."""
        out = np.empty(self.nwalks, dtype=WALK_DTYPE)
        self._check(self._L.gw_walks_read(self._ctx, _ptr(out), self.nwalks), "gw_walks_read")
        return out

    # ------------------------------------------------------------------- run
    def run(self, prog: Program) -> dict:
        st = Stats()
        self._check(self._L.gw_run(self._ctx, C.byref(prog), C.byref(st)), "gw_run")
        return st.as_dict()

    def run_capture(self, prog: Program) -> None:
        """Run prog once, then capture its launch sequence in a hipGraph."""
        self._check(self._L.gw_run_capture(self._ctx, C.byref(prog)), "gw_run_capture")

    def run_replay(self) -> dict:
        """Replay the captured run; stats as run() returns them (step_ms and
        total_ms cover the whole graph)."""
        st = Stats()
        self._check(self._L.gw_run_replay(self._ctx, C.byref(st)), "gw_run_replay")
        return st.as_dict()

    def autotune(self, prog: Program, reps: int = 2) -> tuple[Program, float]:
        """Fastest (block_threads, waves_per_simd) for prog: (program, ms)."""
        best = Program()
        ms = C.c_double()
        self._check(self._L.gw_autotune(self._ctx, C.byref(prog), reps, C.byref(best), C.byref(ms)), "gw_autotune")
        return best, ms.value

    def bucket(self, nblocks: int) -> np.ndarray:
        counts = np.zeros(nblocks, dtype=np.uint64)
        self._check(self._L.gw_bucket(self._ctx, _ptr(counts)), "gw_bucket")
        self.nwalks = int(counts.sum())
        return counts

    def walks_current(self) -> np.ndarray:
        """The loaded walk records in their current slot order."""
        out = np.empty(self.nwalks, dtype=WALK_DTYPE)
        self._check(self._L.gw_walks_current(self._ctx, _ptr(out), self.nwalks), "gw_walks_current")
        return out

    def counts(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.uint32)
        self._check(self._L.gw_counts_read(self._ctx, _ptr(out), self.n), "gw_counts_read")
        return out

    def summary(self) -> tuple[np.ndarray, np.ndarray]:
        h = np.empty(self.nwalks, dtype=np.uint64)
        v = np.empty(self.nwalks, dtype=np.uint32)
        self._check(self._L.gw_summary_read(self._ctx, _ptr(h), _ptr(v), self.nwalks), "gw_summary_read")
        return h, v

    def h2d_probe(self, nbytes: int) -> float:
        ms = C.c_double()
        self._check(self._L.gw_h2d_probe(self._ctx, nbytes, C.byref(ms)), "gw_h2d_probe")
        return ms.value

    def host_gather_probe(self, nbytes: int, gathers: int) -> float:
        """Kernel ms for `gathers` random 4-byte reads of pinned host memory."""
        ms = C.c_double()
        self._check(self._L.gw_host_gather_probe(self._ctx, nbytes, gathers, C.byref(ms)), "gw_host_gather_probe")
        return ms.value

    def copy_probe(self, nbytes: int) -> float:
        ms = C.c_double()
        self._check(self._L.gw_copy_probe(self._ctx, nbytes, C.byref(ms)), "gw_copy_probe")
        return ms.value

    def source_topk(self, k: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Each source's k most-visited vertices (count descending, ties to the
        smaller vertex id), sources ascending: (source, vertex, count)."""
        n = C.c_uint64()
        self._check(self._L.gw_source_topk(self._ctx, k, None, None, None, 0, C.byref(n)), "gw_source_topk")
        s = np.empty(n.value, dtype=np.uint32)
        v = np.empty(n.value, dtype=np.uint32)
        c = np.empty(n.value, dtype=np.uint32)
        self._check(self._L.gw_source_topk(self._ctx, k, _ptr(s), _ptr(v), _ptr(c), n.value, C.byref(n)),
                    "gw_source_topk")
        return s, v, c

    def counts_topk(self, k: int) -> tuple[np.ndarray, np.ndarray]:
        """The k most-visited vertices of the last per-vertex count run (count
        descending, ties to the smaller vertex id): (vertex, count)."""
        m = min(int(k), self.graph_info()["n"]) if k > 0 else 0
        v = np.empty(max(m, 1), dtype=np.uint32)
        c = np.empty(max(m, 1), dtype=np.uint32)
        n = C.c_uint64()
        self._check(self._L.gw_counts_topk(self._ctx, k, _ptr(v), _ptr(c), C.byref(n)), "gw_counts_topk")
        return v[:n.value], c[:n.value]

    def source_counts(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(source, vertex, count) of the last GW_REC_SOURCE_COUNT run, sorted by source then vertex."""
        n = C.c_uint64()
        self._check(self._L.gw_source_counts_read(self._ctx, None, None, None, 0, C.byref(n)), "gw_source_counts_read")
        s = np.empty(n.value, dtype=np.uint32)
        v = np.empty(n.value, dtype=np.uint32)
        c = np.empty(n.value, dtype=np.uint32)
        self._check(self._L.gw_source_counts_read(self._ctx, _ptr(s), _ptr(v), _ptr(c), n.value, C.byref(n)),
                    "gw_source_counts_read")
        return s, v, c

    # --------------------------------------------------------------- serving
    def serve_open(self, prog: Program, pool_walks: int, arena_bytes: int, max_queries: int = 1024,
                   hops_per_step: int = 4) -> None:
        self._check(self._L.gw_serve_open(self._ctx, C.byref(prog), pool_walks, arena_bytes, max_queries,
                                          hops_per_step), "gw_serve_open")
        self._serve_k = 0

    def serve_submit(self, queries) -> None:
        """queries: iterable of (qid, source, walks), or a numpy array of
        QUERY_DTYPE (one call, no per-query Python work)."""
        if isinstance(queries, np.ndarray):
            if queries.dtype != QUERY_DTYPE:
                raise GwError(f"serve_submit: expected dtype {QUERY_DTYPE}, got {queries.dtype}")
            a = np.ascontiguousarray(queries)
            self._check(self._L.gw_serve_submit(self._ctx, _ptr(a), a.shape[0]), "gw_serve_submit")
            return
        q = list(queries)
        arr = (Query * max(1, len(q)))(*[Query(int(a), int(b), int(c)) for a, b, c in q])
        self._check(self._L.gw_serve_submit(self._ctx, arr, len(q)), "gw_serve_submit")

    def serve_step(self) -> dict:
        st = ServeStats()
        self._check(self._L.gw_serve_step(self._ctx, C.byref(st)), "gw_serve_step")
        return st.as_dict()

    def serve_poll(self):
        """Next completed query as (qid, npairs), or None."""
        qid, n = C.c_uint64(), C.c_uint64()
        rc = self._L.gw_serve_poll(self._ctx, C.byref(qid), C.byref(n))
        if rc == 1:
            return None
        self._check(rc, "gw_serve_poll")
        return qid.value, n.value

    def serve_result(self, qid: int, npairs: int, sort: bool = True) -> tuple[np.ndarray, np.ndarray]:
        """A completed query's (vertex, count) pairs (sorted by vertex unless sort=False)."""
        v = np.empty(npairs, dtype=np.uint32)
        c = np.empty(npairs, dtype=np.uint32)
        self._check(self._L.gw_serve_result(self._ctx, qid, _ptr(v), _ptr(c), npairs), "gw_serve_result")
        if sort:
            o = np.argsort(v, kind="stable")
            v, c = v[o], c[o]
        return v, c

    def serve_results(self, max_q: int = 4096, copy: bool = True):
        """Every completed query not yet read, at most max_q, in one call:
        (qids, npairs, vertices, counts) with the queries' pairs back to back
        in completion order.  Needs serve_topk(k > 0) (a query then holds at
        most k pairs); without it, read queries with serve_poll/serve_result.
        copy=False returns views of the engine's buffers, valid until the next
        call (no per-call copies in a serving loop)."""
        k = getattr(self, "_serve_k", 0)
        if k <= 0:
            raise GwError("serve_results needs serve_topk(k > 0)")
        buf = getattr(self, "_serve_buf", None)
        if buf is None or buf[0] < max_q or buf[1][2].shape[0] < max_q * k:
            arrs = (np.empty(max_q, np.uint64), np.empty(max_q, np.uint64),
                    np.empty(max_q * k, np.uint32), np.empty(max_q * k, np.uint32))
            # the ctypes arguments once per buffer, not per call
            buf = (max_q, arrs, tuple(_ptr(a) for a in arrs), C.c_uint32())
            self._serve_buf = buf
        _, (q, npairs, v, c), ptrs, n = buf
        # capacities from the arrays themselves: never more than they hold
        self._check(self._L.gw_serve_results(self._ctx, max_q, *ptrs, v.shape[0], C.byref(n)), "gw_serve_results")
        m = n.value
        tot = int(npairs[:m].sum()) if m else 0
        out = q[:m], npairs[:m], v[:tot], c[:tot]
        return tuple(a.copy() for a in out) if copy else out

    def serve_topk(self, k: int) -> None:
        """Completed queries keep their k most-visited vertices (0 = all pairs)."""
        self._check(self._L.gw_serve_topk(self._ctx, k), "gw_serve_topk")
        self._serve_k = k

    def serve_cache(self, page_log2: int = 0, max_loads: int = 0) -> None:
        """Page size and page copies in flight of a server over a host-resident
        graph (call before serve_open; 0 = defaults)."""
        self._check(self._L.gw_serve_cache(self._ctx, page_log2, max_loads), "gw_serve_cache")

    def serve_close(self) -> None:
        self._check(self._L.gw_serve_close(self._ctx), "gw_serve_close")
        self._serve_k = 0

    def events(self, cap: int) -> np.ndarray:
        """Event log of the last run: (nwalks, cap) u64, (hop << 32) | vertex."""
        out = np.empty((self.nwalks, cap), dtype=np.uint64)
        self._check(self._L.gw_events_read(self._ctx, _ptr(out), self.nwalks, cap), "gw_events_read")
        return out

    def gather_probe(self, nbytes: int, gathers: int) -> float:
        ms = C.c_double()
        self._check(self._L.gw_gather_probe(self._ctx, nbytes, gathers, C.byref(ms)), "gw_gather_probe")
        return ms.value
