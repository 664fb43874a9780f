"""Continuous-batching query serving (gw_serve_*): every query's result is
bit-exact against the oracle, and identical alone and in the mix."""
import numpy as np
import pytest

import graphwalker_amd as gw
import oracle as O

pytestmark = pytest.mark.gpu

SEED = 0x5EED


def expected(off, col, kind, L, qid, source, walks):
    prog = O.program(kind, L, 0.15, off.shape[0] - 1, seed=SEED ^ ((qid >> 32) << 32), stream=qid & 0xFFFFFFFF)
    c = O.replay(prog, off, col, O.make_walks(np.full(walks, source, np.uint32)))["counts"]
    nz = np.nonzero(c)[0]
    return nz.astype(np.uint32), c[nz]


def serve_all(engine, arrivals, kind, L, hops, pool, arena, max_queries=64, topk=0, batch=0):
    """arrivals: list of (step, qid, source, walks); submits each at its step.
    batch > 0 reads the results through serve_results, at most batch queries
    a call; otherwise one query at a time through serve_poll / serve_result."""
    engine.serve_open(gw.program(kind, L, 0.15, 0, seed=SEED), pool, arena, max_queries, hops)
    if topk:
        engine.serve_topk(topk)
    results, step, pending = {}, 0, sorted(arrivals)
    try:
        while pending or len(results) < len(arrivals):
            now = [a for a in pending if a[0] <= step]
            pending = [a for a in pending if a[0] > step]
            if now:
                engine.serve_submit([(q, s, w) for _, q, s, w in now])
            st = engine.serve_step()
            assert st["live_walks"] <= pool
            if batch:
                while (got := engine.serve_results(batch))[0].shape[0]:
                    q, npairs, v, c = got
                    assert q.shape[0] <= batch and int(npairs.sum()) == v.shape[0] == c.shape[0]
                    ends = np.cumsum(npairs.astype(np.int64)).tolist()
                    for qid, e, m in zip(q.tolist(), ends, npairs.tolist()):
                        results[qid] = v[e - m:e], c[e - m:e]
            else:
                while (r := engine.serve_poll()) is not None:
                    results[r[0]] = engine.serve_result(*r, sort=not topk)
            step += 1
            assert step < 10_000
    finally:
        engine.serve_close()
    return results


@pytest.fixture(scope="module")
def g12(engine):
    off, col = O.rmat_graph(12, 1 << 16, seed=7)
    engine.graph_set(off, col)
    return off, col


@pytest.mark.parametrize("kind,hops", [(gw.RESTART, 3), (gw.STOP, 2), (gw.RESTART, 10)])
def test_served_results_bit_exact(engine, g12, kind, hops):
    off, col = g12
    rng = np.random.default_rng(kind * 10 + hops)
    arrivals = [(int(rng.integers(0, 8)), (int(rng.integers(0, 4)) << 32) | i, int(rng.integers(0, 4096)),
                 int(rng.integers(50, 3000))) for i in range(40)]
    res = serve_all(engine, arrivals, kind, 10, hops, pool=200_000, arena=64 << 20)
    assert len(res) == len(arrivals)
    for _, qid, src, w in arrivals:
        v, c = res[qid]
        ev, ec = expected(off, col, kind, 10, qid, src, w)
        assert np.array_equal(v, ev) and np.array_equal(c, ec), qid


def test_result_alone_equals_result_in_mix(engine, g12):
    rng = np.random.default_rng(5)
    arrivals = [(i // 5, 1000 + i, int(rng.integers(0, 4096)), int(rng.integers(100, 5000))) for i in range(25)]
    mixed = serve_all(engine, arrivals, gw.RESTART, 10, 3, pool=60_000, arena=64 << 20)
    for a in arrivals[::6]:
        alone = serve_all(engine, [(0,) + a[1:]], gw.RESTART, 10, 3, pool=60_000, arena=64 << 20)
        assert np.array_equal(alone[a[1]][0], mixed[a[1]][0])
        assert np.array_equal(alone[a[1]][1], mixed[a[1]][1])


def test_admission_is_bounded_by_the_pool(engine, g12):
    # 30 queries of 3000 walks with a 7000-walk pool: at most two in flight
    arrivals = [(0, i, i * 37 % 4096, 3000) for i in range(30)]
    res = serve_all(engine, arrivals, gw.RESTART, 10, 4, pool=7000, arena=64 << 20)
    assert sorted(res) == list(range(30))
    off, col = g12
    v, c = res[17]
    ev, ec = expected(off, col, gw.RESTART, 10, 17, 17 * 37 % 4096, 3000)
    assert np.array_equal(v, ev) and np.array_equal(c, ec)


def test_serve_rejects_bad_queries(engine, g12):
    engine.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=1), 1000, 1 << 20, 8, 4)
    try:
        with pytest.raises(gw.GwError):
            engine.serve_submit([(1, 1 << 20, 10)])
        with pytest.raises(gw.GwError):
            engine.serve_submit([(1, 5, 5000)])  # more walks than the pool
    finally:
        engine.serve_close()
    # a query whose count region can never fit the arena is refused at once
    # (it would otherwise block the queue), and so is an id already in use
    engine.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=1), 100_000, 32 << 10, 8, 4)
    try:
        with pytest.raises(gw.GwError, match="whole arena"):
            engine.serve_submit([(1, 5, 5000)])
        engine.serve_submit([(2, 5, 100)])
        with pytest.raises(gw.GwError, match="already"):
            engine.serve_submit([(2, 6, 100)])
        with pytest.raises(gw.GwError, match="already"):
            engine.serve_submit([(3, 6, 100), (3, 7, 100)])
        while (r := engine.serve_poll()) is None:
            engine.serve_step()
        assert r[0] == 2
        with pytest.raises(gw.GwError, match="already"):
            engine.serve_submit([(2, 6, 100)])  # completed, not yet read
        engine.serve_result(*r)
        engine.serve_submit([(2, 6, 100)])      # read: the id is free again
    finally:
        engine.serve_close()
    with pytest.raises(gw.GwError, match="at most 65535"):
        engine.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=1), 1000, 1 << 20, 70_000, 4)


@pytest.mark.parametrize("k,batch", [(1, 0), (10, 0), (1024, 0), (10, 3), (1024, 64)])
def test_served_topk(engine, g12, k, batch):
    # completed queries keep their k most-visited vertices, count descending,
    # ties to the smaller vertex -- the top of the exact counts; read one by
    # one or in batches of at most `batch` queries
    off, col = g12
    rng = np.random.default_rng(k)
    arrivals = [(int(rng.integers(0, 5)), 500 + i, int(rng.integers(0, 4096)), int(rng.integers(100, 4000)))
                for i in range(20)]
    res = serve_all(engine, arrivals, gw.RESTART, 10, 3, pool=100_000, arena=64 << 20, topk=k, batch=batch)
    assert len(res) == len(arrivals)
    for _, qid, src, w in arrivals:
        v, c = res[qid]
        ev, ec = expected(off, col, gw.RESTART, 10, qid, src, w)
        order = np.lexsort((ev, -ec.astype(np.int64)))[:k]
        assert np.array_equal(v, ev[order]) and np.array_equal(c, ec[order]), qid


@pytest.mark.parametrize("k", [100, 1024])
def test_served_topk_many_chunks(engine, k):
    # queries with more pairs than one top-k chunk (32768 items): the chunks'
    # candidates merge to the exact top-k of the whole query
    off, col = O.rmat_graph(17, 1 << 21, seed=9)
    engine.graph_set(off, col)
    arrivals = [(0, 700, 11, 60_000), (0, 701, 5000, 40_000), (1, 702, 77, 2000), (1, 703, 123_456, 25_000)]
    res = serve_all(engine, arrivals, gw.RESTART, 10, 10, pool=300_000, arena=256 << 20, topk=k)
    most = 0
    for _, qid, src, w in arrivals:
        v, c = res[qid]
        ev, ec = expected(off, col, gw.RESTART, 10, qid, src, w)
        most = max(most, ev.shape[0])
        order = np.lexsort((ev, -ec.astype(np.int64)))[:k]
        assert np.array_equal(v, ev[order]) and np.array_equal(c, ec[order]), qid
    assert most > 32768


def test_batched_results_skip_queries_read_by_id(engine, g12):
    # a query read through serve_result by its id before any poll is not
    # handed out again by serve_poll or serve_results; serve_results needs a
    # top-k (its output holds at most k pairs a query)
    off, col = g12
    engine.graph_set(off, col)  # an earlier test may have set another graph
    engine.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=SEED), 100_000, 64 << 20, 64, 10)
    try:
        with pytest.raises(gw.GwError, match="serve_topk"):
            engine.serve_results()
        engine.serve_topk(5)
        engine.serve_submit([(1, 10, 300), (2, 20, 300), (3, 30, 300)])
        for _ in range(100):
            engine.serve_step()
        v, c = engine.serve_result(2, 5, sort=False)
        ev, ec = expected(off, col, gw.RESTART, 10, 2, 20, 300)
        order = np.lexsort((ev, -ec.astype(np.int64)))[:5]
        assert np.array_equal(v, ev[order]) and np.array_equal(c, ec[order])
        q, npairs, v, c = engine.serve_results(10)
        assert sorted(q.tolist()) == [1, 3] and list(npairs) == [5, 5]
        assert engine.serve_poll() is None and engine.serve_results()[0].shape[0] == 0
    finally:
        engine.serve_close()


def test_submit_numpy_queries(engine, g12):
    # queries handed over as one QUERY_DTYPE array are served like tuples
    off, col = g12
    engine.graph_set(off, col)  # an earlier test may have set another graph
    engine.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=SEED), 100_000, 64 << 20, 64, 10)
    try:
        qids, srcs, walks = [7, 8, 9], [10, 20, 30], [200, 300, 400]
        q = np.zeros(3, gw.QUERY_DTYPE)
        q["qid"], q["source"], q["walks"] = qids, srcs, walks
        engine.serve_submit(q)
        for _ in range(50):
            engine.serve_step()
        got = {}
        while (r := engine.serve_poll()) is not None:
            got[r[0]] = engine.serve_result(*r)
        assert sorted(got) == qids
        for qid, src, w in zip(qids, srcs, walks):
            ev, ec = expected(off, col, gw.RESTART, 10, qid, src, w)
            assert np.array_equal(got[qid][0], ev) and np.array_equal(got[qid][1], ec)
        with pytest.raises(gw.GwError, match="dtype"):
            engine.serve_submit(np.zeros(2, np.uint64))
    finally:
        engine.serve_close()


def test_batched_results_capacity(engine, g12):
    # gw_serve_results stops before a query whose pairs would pass `cap`, and
    # refuses a call whose first query does not fit; views (copy=False) hold
    # the same pairs as copies
    import ctypes as C
    off, col = g12
    engine.graph_set(off, col)  # an earlier test may have set another graph
    engine.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=SEED), 100_000, 64 << 20, 64, 10)
    try:
        engine.serve_topk(8)
        # two sources with out-edges: each query then has more than 8 distinct vertices
        s41, s42 = (int(x) for x in np.nonzero(np.diff(off) > 2)[0][:2])
        srcs = {41: s41, 42: s42}
        engine.serve_submit([(41, s41, 300), (42, s42, 300)])
        engine.serve_step()
        L, ctx = engine._L, engine._ctx
        q = np.zeros(4, np.uint64)
        npairs = np.zeros(4, np.uint64)
        v = np.zeros(16, np.uint32)
        c = np.zeros(16, np.uint32)
        n = C.c_uint32()
        p = gw._ptr
        assert L.gw_serve_results(ctx, 4, p(q), p(npairs), p(v), p(c), 4, C.byref(n)) == -1  # GW_EINVAL
        assert L.gw_serve_results(ctx, 4, p(q), p(npairs), p(v), p(c), 12, C.byref(n)) == 0
        assert n.value == 1 and npairs[0] == 8  # the second query's 8 pairs would pass 12
        first = q[0]
        views = engine.serve_results(4, copy=False)
        assert views[0].shape[0] == 1 and views[0][0] != first and views[1][0] == 8
        for qid, src, vv, cc in ((first, srcs[int(first)], v[:8], c[:8]),
                                 (views[0][0], srcs[int(views[0][0])], views[2], views[3])):
            ev, ec = expected(off, col, gw.RESTART, 10, int(qid), src, 300)
            order = np.lexsort((ev, -ec.astype(np.int64)))[:8]
            assert np.array_equal(vv, ev[order]) and np.array_equal(cc, ec[order])
    finally:
        engine.serve_close()
