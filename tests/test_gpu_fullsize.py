"""The configured workloads at their full sizes, checked through properties
that hold at any size: a well-formed device-built CSR, exact step totals,
identical results across schedules and counting methods, a sample of walks
(or sources, or pairs) bit-exact against the oracle on the same graph, and
transposition applied twice giving back the graph.  Each case names its
sizes; the oracle replays only the sample."""
import numpy as np
import pytest

import graphwalker_amd as gw
import oracle as O

pytestmark = pytest.mark.gpu


def oprog(p: gw.Program) -> O.Program:
    return O.Program(p.kind, p.L, p.stop_thr, p.nvertices, p.stream, p.seed)


def _csr_ok(off, col, n, m):
    assert off.shape == (n + 1,) and col.shape == (m,)
    assert int(off[0]) == 0 and int(off[-1]) == m
    assert np.all(off[1:] >= off[:-1])
    assert int(col.max()) < n


def test_rawwalks_rmat24_fullsize(engine):
    """R-MAT scale 24, edge factor 16, seed 2; one walk per vertex, jump with
    p = 0, L = 10: 167,772,160 steps whatever the schedule."""
    n = 1 << 24
    m = 16 * n
    engine.graph_rmat(24, m, 2)
    off, col = engine.graph_read()
    _csr_ok(off, col, n, m)
    engine.walks_init_per_vertex(1)
    out = {}
    for sched, layout in (("persistent", "ext"), ("plain", "csr"), ("blocked", "ext")):
        if sched == "blocked":
            engine.graph_blocks(np.arange(0, n + 1, 1 << 20, dtype=np.uint32))
        prog = gw.program(gw.JUMP, 10, 0.0, n, seed=2, recorder=gw.REC_VERTEX_COUNT, schedule=sched, layout=layout)
        st = engine.run(prog)
        assert st["steps"] == 10 * n and st["visits"] == 10 * n, sched
        c = engine.counts()
        assert int(c.sum(dtype=np.uint64)) == 10 * n
        out[sched] = (prog, c.copy(), engine.walks_final())
    prog, c, fin = out["persistent"]
    for other in ("plain", "blocked"):
        assert np.array_equal(out[other][1], c) and np.array_equal(out[other][2], fin), other
    assert np.array_equal(fin["id"], np.arange(n, dtype=np.uint32)) and np.all(fin["hop"] == 10)
    ids = np.sort(np.random.default_rng(1).choice(n, 4096, replace=False)).astype(np.uint32)
    ref = O.replay(oprog(prog), off, col, O.make_walks(ids, sources_=ids, ids=ids), counts=False, per_walk=True)
    assert np.array_equal(fin[ids], ref["final"])


def test_ppr4096_rmat26_fullsize(engine):
    """R-MAT scale 26, edge factor 16, seed 3; 4096 uniform sources x 2000
    stop walks (p 0.15, cap 20): per-source counts identical by the
    partitioned and the global sort-reduce methods, the counts summing to
    the run's visits, sampled sources holding at least one visit of their
    own vertex per walk, and 6 sources replayed by the oracle."""
    n = 1 << 26
    m = 16 * n
    engine.graph_rmat(26, m, 3)
    S, W, L = 4096, 2000, 20
    src = gw.sources(3, n, S)
    engine.walks_init_sources(src, W, source_field=np.arange(S, dtype=np.uint32))
    res = {}
    for method in (gw.SRC_PARTITIONED, gw.SRC_GLOBAL):
        prog = gw.program(gw.STOP, L, 0.15, n, seed=3, recorder=gw.REC_SOURCE_COUNT | gw.REC_NO_FINAL,
                          source_method=method)
        st = engine.run(prog)
        assert st["source_method"] == method
        s, v, c = engine.source_counts()
        assert int(c.sum(dtype=np.uint64)) == st["visits"]
        assert np.all(np.diff(s.astype(np.int64)) >= 0)
        res[method] = (st, s, v, c)
    a, b = res[gw.SRC_PARTITIONED], res[gw.SRC_GLOBAL]
    assert a[0]["steps"] == b[0]["steps"] and a[0]["visits"] == b[0]["visits"]
    for i in (1, 2, 3):
        assert np.array_equal(a[i], b[i])
    s, v, c = a[1], a[2], a[3]
    bounds = np.searchsorted(s, np.arange(S + 1))
    # every walk visits its source first
    first = np.array([c[bounds[k]:bounds[k + 1]][v[bounds[k]:bounds[k + 1]] == src[k]][0] for k in range(0, S, 97)])
    assert np.all(first >= W)
    off, col = engine.graph_read()
    _csr_ok(off, col, n, m)
    for k in (0, 1, 777, 2048, 4000, 4095):
        w = O.make_walks(np.full(W, src[k], np.uint32), sources_=np.full(W, k, np.uint32),
                         ids=k * W + np.arange(W, dtype=np.uint32))
        cnt = O.replay(oprog(prog), off, col, w)["counts"]
        nz = np.nonzero(cnt)[0]
        lo, hi = bounds[k], bounds[k + 1]
        assert np.array_equal(v[lo:hi], nz.astype(np.uint32)) and np.array_equal(c[lo:hi], cnt[nz]), k


def _transposed_np(off, col):
    n = off.shape[0] - 1
    toff = np.zeros(n + 1, dtype=np.uint64)
    toff[1:] = np.cumsum(np.bincount(col, minlength=n).astype(np.uint64))
    return toff


def test_simrank_rmat28_fullsize(engine):
    """R-MAT scale 28, edge factor 16, seed 4 (2^32 edges): the device
    transpose has the in-degree offsets, transposing twice gives back the
    graph, and 2 pairs x 1000 reverse walks of length 11 give the oracle's
    first-meeting counts on the transposed graph."""
    n = 1 << 28
    m = 16 * n
    engine.graph_rmat(28, m, 4)
    off, col = engine.graph_read()
    _csr_ok(off, col, n, m)
    engine.graph_transpose()
    toff, tcol = engine.graph_read()
    _csr_ok(toff, tcol, n, m)
    assert np.array_equal(toff, _transposed_np(off, col))
    del toff
    engine.graph_transpose()
    roff, rcol = engine.graph_read()
    assert np.array_equal(roff, off) and np.array_equal(rcol, col)
    del roff, rcol, off, col
    engine.graph_transpose()
    toff, tcol = engine.graph_read()
    R, L = 1000, 11
    pairs = gw.sources(4, n, 4).reshape(2, 2)
    a, b = pairs[:, 0].copy(), pairs[:, 1].copy()
    engine.walks_init_pairs(a, b, R, L)
    prog = gw.program(gw.RESTART_JOINT, L, 0.15, n, seed=4, recorder=gw.REC_EVENTS, event_cap=L - 1)
    st = engine.run(prog)
    assert st["steps"] == 2 * 2 * R * (L - 1)
    got = engine.meeting_counts(2, R, L)
    op = oprog(prog)
    for k in range(2):
        paths = []
        for side, start in ((0, a[k]), (1, b[k])):
            P = np.full((R, L - 1), -1, dtype=np.int64)
            for i in range(R):
                w = O.make_walks([start], sources_=[start], hops=[i * L], ids=[2 * k * R + side * R + i])[0]
                tv, th, _ = O.trace(op, toff, tcol, w)
                P[i, (th - i * L).astype(np.int64)] = tv
            paths.append(P)
        eq = (paths[0][:, None, :] == paths[1][None, :, :]) & (paths[0][:, None, :] >= 0)
        hit = eq.any(axis=2)
        exp = np.zeros(L, dtype=np.uint64)
        np.add.at(exp, eq.argmax(axis=2)[hit], 1)
        assert np.array_equal(got[k].astype(np.uint64), exp), k
