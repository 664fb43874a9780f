"""gw_run_noalloc (the allocation-free run) against gw_run, through the C ABI.

Device buffers for its inputs and outputs come from the HIP runtime the
library itself links (hipMalloc / hipMemcpy through ctypes), so the only
thing under test is the library.  Two calls in a row, the second on a
smaller input, must each equal gw_run's persistent-schedule results on the
same walks bit for bit: the static scratch keeps the last call's counters,
so a missing clear would show up in the second call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import graphwalker_amd as gw

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")  # already loaded by libgwhip: the same runtime
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def malloc(self, nbytes: int) -> C.c_void_p:
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(1, nbytes)) == 0, "hipMalloc"
        self.bufs.append(p)
        return p

    def h2d(self, dst, a: np.ndarray) -> None:
        assert self.L.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0, "hipMemcpy H2D"

    def d2h(self, a: np.ndarray, src) -> None:
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2) == 0, "hipMemcpy D2H"

    def sync(self) -> None:
        assert self.L.hipDeviceSynchronize() == 0, "hipDeviceSynchronize"

    def free_all(self) -> None:
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(eng, hip, prog, walks: np.ndarray, with_counts: bool):
    lib = gw.lib()
    d_w = hip.malloc(walks.nbytes)
    hip.h2d(d_w, walks)
    d_c = hip.malloc(eng.n * 4) if with_counts else None
    d_t = hip.malloc(24)
    rc = lib.gw_run_noalloc(eng._ctx, C.byref(prog), d_w, walks.shape[0], d_c, d_t, None)
    assert rc == 0, lib.gw_last_error(eng._ctx)
    hip.sync()
    tot = np.zeros(3, np.uint64)
    hip.d2h(tot, d_t)
    counts = None
    if with_counts:
        counts = np.zeros(eng.n, np.uint32)
        hip.d2h(counts, d_c)
    return tot, counts


def _reference(eng, prog, walks: np.ndarray):
    eng.walks_load(walks)
    st = eng.run(prog)
    return st, (eng.counts() if prog.recorder & gw.REC_VERTEX_COUNT else None)


@pytest.mark.parametrize("kind,layout", [(gw.STOP, gw.LAYOUT_CSR), (gw.RESTART, gw.LAYOUT_EXT), (gw.JUMP, gw.LAYOUT_CSR)])
def test_noalloc_equals_run_twice(engine, kind, layout):
    eng = engine
    eng.graph_rmat(14, 1 << 18, 21)
    n = eng.n
    hip = _Hip()
    try:
        srcs = gw.sources(5, n, 64)
        big = np.zeros(64 * 700, dtype=gw.WALK_DTYPE)
        big["id"] = np.arange(big.shape[0])
        big["source"] = np.repeat(srcs, 700)
        big["vertex"] = big["source"]
        small = big[: 64 * 90].copy()
        small["hop"] = 3  # walks that start mid-way
        for rec in (gw.REC_VERTEX_COUNT | gw.REC_NO_FINAL, gw.REC_NO_FINAL):
            prog = gw.program(kind, 12, 0.15, n, seed=9, recorder=rec, schedule=gw.SCHED_PERSISTENT, layout=layout)
            for walks in (big, small):  # two calls in a row, the second smaller
                st, ref_counts = _reference(eng, prog, walks)
                tot, counts = _noalloc(eng, hip, prog, walks, (rec & gw.REC_VERTEX_COUNT) != 0)
                assert int(tot[0]) == st["steps"] and int(tot[1]) == st["visits"], (tot, st)
                assert int(tot[2]) == 0
                if ref_counts is not None:
                    assert np.array_equal(counts, ref_counts)
    finally:
        hip.free_all()


def test_noalloc_skips_walks_outside_the_graph(engine):
    eng = engine
    eng.graph_rmat(12, 1 << 15, 3)
    n = eng.n
    hip = _Hip()
    try:
        good = np.zeros(5000, dtype=gw.WALK_DTYPE)
        good["id"] = np.arange(5000)
        good["source"] = good["vertex"] = np.arange(5000) % n
        prog = gw.program(gw.RESTART, 10, 0.15, n, seed=4, recorder=gw.REC_VERTEX_COUNT | gw.REC_NO_FINAL,
                          schedule=gw.SCHED_PERSISTENT, layout=gw.LAYOUT_CSR)
        st, ref_counts = _reference(eng, prog, good)
        bad = np.concatenate([good[:100], good[:3], good[100:]])
        bad["vertex"][100] = n          # vertex outside the graph
        bad["vertex"][101] = 0xFFFFFFFF
        bad["source"][102] = n + 7      # restart target outside the graph
        bad["id"][100:103] = 999999
        tot, counts = _noalloc(eng, hip, prog, bad, True)
        assert int(tot[2]) == 3
        assert int(tot[0]) == st["steps"] and int(tot[1]) == st["visits"]
        assert np.array_equal(counts, ref_counts)
    finally:
        hip.free_all()


def test_noalloc_rejects(engine):
    eng = engine
    eng.graph_rmat(10, 1 << 12, 1)
    lib = gw.lib()
    t = C.c_void_p(1)  # never dereferenced: every call below fails its host checks
    prog = gw.program(gw.STOP, 10, 0.15, eng.n, recorder=gw.REC_WALK_SUMMARY)
    assert lib.gw_run_noalloc(eng._ctx, C.byref(prog), t, 10, None, t, None) == -1
    prog = gw.program(gw.STOP, 10, 0.15, eng.n, recorder=gw.REC_NONE)
    assert lib.gw_run_noalloc(eng._ctx, C.byref(prog), t, 1 << 32, None, t, None) == -1
    assert b"2^32" in lib.gw_last_error(eng._ctx)
    prog = gw.program(gw.STOP, 10, 0.15, eng.n, recorder=gw.REC_VERTEX_COUNT)
    assert lib.gw_run_noalloc(eng._ctx, C.byref(prog), t, 10, None, t, None) == -1
    prog = gw.program(gw.STOP, 10, 0.15, eng.n, schedule=gw.SCHED_BLOCKED)
    assert lib.gw_run_noalloc(eng._ctx, C.byref(prog), t, 10, None, t, None) == -1
