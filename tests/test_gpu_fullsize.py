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
    # FP64 on the configuration's own graph: the expected visit vectors of the
    # first 8 sources with out-edges (20-hop stop chain, pulled over the
    # device-transposed CSR).  Per source: the normalised counts within the
    # Monte-Carlo L1 bound of its 2000 walks, the visit total, and each of the
    # 50 most-visited vertices within 5 sigma (a vertex's count sums 2000
    # per-walk counts in [0, L], so its variance is at most L x expected)
    from oracle import fp64
    engine.graph_transpose()
    toff, tcol = engine.graph_read()
    deg = off[src.astype(np.int64) + 1].astype(np.int64) - off[src.astype(np.int64)].astype(np.int64)
    ks = [k for k in range(S) if deg[k] > 0][:8]
    assert len(ks) == 8
    E = O.fp64_stop_visits(off, toff, tcol, src[ks], float(W), L, prog.stop_thr / 2.0 ** 32)
    del toff, tcol
    for i, k in enumerate(ks):
        cnt = np.zeros(n, dtype=np.float64)
        lo, hi = bounds[k], bounds[k + 1]
        cnt[v[lo:hi]] = c[lo:hi]
        exp = E[i]
        assert abs(cnt.sum() - exp.sum()) <= 5 * np.sqrt(L * exp.sum()), k
        l1, bound = fp64.l1_normalised(cnt, exp), fp64.mc_l1_bound(exp, W, exp.sum() / W, L)
        top = np.argsort(exp)[-50:]
        z = np.abs(cnt[top] - exp[top]) / np.sqrt(L * exp[top])
        print(f"[ppr fp64] source {k} (vertex {src[k]}, out-degree {deg[k]}): L1 {l1:.4f} (MC bound {bound:.4f}), "
              f"top-50 vertices max |z| {z.max():.2f}, own vertex {cnt[src[k]]:.0f} vs {exp[src[k]]:.1f}")
        assert l1 < bound, (k, l1, bound)
        assert np.all(z <= 5.0), (k, z.max())


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
    # FP64 SimRank on a sampled subgraph of this graph: the vertices within
    # two in-link hops of 4 uniform pairs (at most 10 in-neighbours taken per
    # vertex) and every edge among them; the device's first-meeting rates of
    # 8 copies x 1024 walks per endpoint against the exact pair chain there
    from oracle import fp64
    cand = gw.sources(44, n, 4096)
    ends = cand[(toff[cand.astype(np.int64) + 1] - toff[cand.astype(np.int64)]) >= 3][:8]  # endpoints with in-links
    assert ends.shape[0] == 8
    ball, frontier = set(int(x) for x in ends), [int(x) for x in ends]
    for _ in range(2):
        nxt = []
        for u in frontier:
            for w in tcol[int(toff[u]):min(int(toff[u + 1]), int(toff[u]) + 10)].tolist():
                if w not in ball:
                    ball.add(w)
                    nxt.append(w)
        frontier = nxt
    S = np.array(sorted(ball), dtype=np.uint32)
    sub_src, sub_dst = [], []
    for i, u in enumerate(S.tolist()):
        r = tcol[int(toff[u]):int(toff[u + 1])]
        r = r[np.isin(r, S)]
        sub_dst.append(np.searchsorted(S, r).astype(np.uint32))
        sub_src.append(np.full(r.shape[0], i, np.uint32))
    del tcol
    local = {int(u): i for i, u in enumerate(S.tolist())}
    soff, scol = O.csr_build(S.shape[0], np.concatenate(sub_src), np.concatenate(sub_dst))
    engine.graph_set(soff, scol)
    R2, copies = 1024, 8
    la = np.repeat(np.array([local[int(x)] for x in ends[0::2]], np.uint32), copies)
    lb = np.repeat(np.array([local[int(x)] for x in ends[1::2]], np.uint32), copies)
    engine.walks_init_pairs(la, lb, R2, L)
    sprog = gw.program(gw.RESTART_JOINT, L, 0.15, S.shape[0], seed=41, recorder=gw.REC_EVENTS, event_cap=L - 1)
    engine.run(sprog)
    rates = engine.meeting_counts(la.shape[0], R2, L).astype(np.float64) / (R2 * R2)
    pr = sprog.stop_thr / 2.0 ** 32
    for j in range(4):
        f = fp64.first_meeting(soff, scol, int(la[j * copies]), int(lb[j * copies]), pr, L - 1)
        est = rates[j * copies:(j + 1) * copies, :L - 1].mean(axis=0)
        tol = 5.0 * np.sqrt((f / (R2 * R2) + 2.0 * f / R2) / copies) + 1e-9
        print(f"[simrank fp64] subgraph of {S.shape[0]} vertices, pair {j}: SimRank estimate "
              f"{float(est @ 0.8 ** np.arange(L - 1)):.5f}, exact {float(f @ 0.8 ** np.arange(L - 1)):.5f}")
        assert np.all(np.abs(est - f) <= tol), (j, est, f, tol)
