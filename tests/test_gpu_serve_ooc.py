"""Serving over host-resident graphs, and the host-resident R-MAT build.

The server over a host-resident graph reads columns through a device page
cache smaller than the graph (the rest straight from pinned host memory), so
these tests give it caches that hold some pages and force loads and
evictions: every served result must still equal the oracle's replay bit for
bit, and a query alone must equal the same query in the mix.  The host
R-MAT build, cut into many source ranges by a small device budget, must equal
the device build and the oracle's graph bit for bit.
"""
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


def serve_all(engine, arrivals, kind, L, hops, pool, arena, page_log2=0, max_loads=0, topk=0):
    """arrivals: (step, qid, source, walks); returns results and the summed page statistics."""
    engine.serve_cache(page_log2, max_loads)
    engine.serve_open(gw.program(kind, L, 0.15, 0, seed=SEED, layout=gw.LAYOUT_AUTO), pool, arena, 64, hops)
    if topk:
        engine.serve_topk(topk)
    results, step, pending = {}, 0, sorted(arrivals)
    tot = {"pages_loaded": 0, "sampled_reads": 0, "sampled_resident": 0, "h2d_bytes": 0, "pages_resident": 0}
    try:
        while pending or len(results) < len(arrivals):
            now = [a for a in pending if a[0] <= step]
            pending = [a for a in pending if a[0] > step]
            if now:
                engine.serve_submit([(q, s, w) for _, q, s, w in now])
            st = engine.serve_step()
            for k in ("pages_loaded", "sampled_reads", "sampled_resident", "h2d_bytes"):
                tot[k] += st[k]
            tot["pages_resident"] = max(tot["pages_resident"], st["pages_resident"])
            while (r := engine.serve_poll()) is not None:
                results[r[0]] = engine.serve_result(*r, sort=not topk)
            step += 1
            assert step < 10_000
    finally:
        engine.serve_close()
        engine.serve_cache(0, 0)
    return results, tot


@pytest.fixture(scope="module")
def g12():
    return O.rmat_graph(12, 1 << 16, seed=7)


def host_graph(engine, off, col, resident_pages, page_log2=10):
    """Host-resident graph whose page cache holds the offsets and `resident_pages` pages."""
    cache = (off.shape[0]) * 8 + resident_pages * (4 << page_log2)
    engine.graph_set_host(off, col, cache)


@pytest.mark.parametrize("kind,hops,resident", [(gw.RESTART, 3, 20), (gw.STOP, 2, 20), (gw.RESTART, 10, 5),
                                                (gw.RESTART, 4, 0), (gw.RESTART, 3, 64)])
def test_served_over_host_graph_bit_exact(engine, g12, kind, hops, resident):
    off, col = g12
    host_graph(engine, off, col, resident)
    rng = np.random.default_rng(kind * 10 + hops + resident)
    arrivals = [(int(rng.integers(0, 12)), (int(rng.integers(0, 4)) << 32) | i, int(rng.integers(0, 4096)),
                 int(rng.integers(50, 3000))) for i in range(40)]
    res, tot = serve_all(engine, arrivals, kind, 10, hops, pool=200_000, arena=64 << 20, page_log2=10, max_loads=2)
    assert len(res) == len(arrivals)
    for _, qid, src, w in arrivals:
        v, c = res[qid]
        ev, ec = expected(off, col, kind, 10, qid, src, w)
        assert np.array_equal(v, ev) and np.array_equal(c, ec), qid
    assert tot["sampled_reads"] > 0
    assert tot["pages_resident"] <= resident
    if resident == 0:
        assert tot["sampled_resident"] == 0 and tot["pages_loaded"] == 0
    elif resident < 64:
        # demand moved pages in (64 pages of 1024 edges, the cache holds fewer)
        assert tot["pages_loaded"] > 0 and 0 < tot["sampled_resident"] < tot["sampled_reads"]
    else:
        assert tot["sampled_resident"] == tot["sampled_reads"]


def test_host_graph_alone_equals_mix(engine, g12):
    off, col = g12
    host_graph(engine, off, col, 12)
    rng = np.random.default_rng(11)
    arrivals = [(i // 5, 3000 + i, int(rng.integers(0, 4096)), int(rng.integers(100, 5000))) for i in range(25)]
    mixed, _ = serve_all(engine, arrivals, gw.RESTART, 10, 3, pool=60_000, arena=64 << 20, page_log2=10, max_loads=3)
    for a in arrivals[::6]:
        alone, _ = serve_all(engine, [(0,) + a[1:]], gw.RESTART, 10, 3, pool=60_000, arena=64 << 20, page_log2=10)
        assert np.array_equal(alone[a[1]][0], mixed[a[1]][0])
        assert np.array_equal(alone[a[1]][1], mixed[a[1]][1])


def test_host_graph_served_topk(engine, g12):
    off, col = g12
    host_graph(engine, off, col, 16)
    rng = np.random.default_rng(3)
    arrivals = [(int(rng.integers(0, 5)), 900 + i, int(rng.integers(0, 4096)), int(rng.integers(100, 4000)))
                for i in range(12)]
    res, _ = serve_all(engine, arrivals, gw.RESTART, 10, 3, pool=100_000, arena=64 << 20, page_log2=10, topk=50)
    for _, qid, src, w in arrivals:
        v, c = res[qid]
        ev, ec = expected(off, col, gw.RESTART, 10, qid, src, w)
        order = np.lexsort((ev, -ec.astype(np.int64)))[:50]
        assert np.array_equal(v, ev[order]) and np.array_equal(c, ec[order]), qid


def test_host_graph_server_refusals(engine, g12):
    off, col = g12
    engine.graph_set_host(off, col, 1024)  # smaller than the offsets
    with pytest.raises(gw.GwError, match="offsets"):
        engine.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=1), 1000, 1 << 20, 8, 4)
    host_graph(engine, off, col, 4)
    with pytest.raises(gw.GwError, match="CSR layout"):
        engine.serve_open(gw.program(gw.RESTART, 10, 0.15, 0, seed=1, layout=gw.LAYOUT_EXT), 1000, 1 << 20, 8, 4)
    with pytest.raises(gw.GwError):
        engine.serve_cache(9, 0)


@pytest.mark.parametrize("scale,work", [(12, 0), (16, 8 << 20), (20, 96 << 20)])
def test_rmat_host_equals_device_build(engine, scale, work):
    m = 16 << scale
    engine.graph_rmat(scale, m, 5)
    doff, dcol = engine.graph_read()
    ms = engine.graph_rmat_host(scale, m, 5, cache_bytes=1 << 30, work_bytes=work)
    assert ms > 0
    hoff, hcol = engine.graph_read()
    assert np.array_equal(hoff, doff) and np.array_equal(hcol, dcol)
    if scale <= 16:
        roff, rcol = O.rmat_graph(scale, m, 5)
        assert np.array_equal(hoff, roff) and np.array_equal(hcol, rcol)


def test_rmat_host_refuses_a_budget_below_one_bin(engine):
    with pytest.raises(gw.GwError, match="device memory"):
        engine.graph_rmat_host(16, 1 << 20, 5, cache_bytes=1 << 30, work_bytes=64 << 10)


def test_served_over_host_rmat_equals_device_graph(engine):
    # the same queries served over the device-resident graph and over the
    # host-built, host-resident one (a quarter of its pages cached)
    scale, m = 14, 16 << 14
    rng = np.random.default_rng(2)
    arrivals = [(int(rng.integers(0, 6)), 77 + i, int(rng.integers(0, 1 << scale)), int(rng.integers(200, 6000)))
                for i in range(30)]
    engine.graph_rmat(scale, m, 9)
    dev, _ = serve_all(engine, arrivals, gw.RESTART, 10, 10, pool=200_000, arena=64 << 20)
    npages = m >> 12
    engine.graph_rmat_host(scale, m, 9, cache_bytes=((1 << scale) + 1) * 8 + (npages // 4) * (4 << 12), work_bytes=8 << 20)
    host, tot = serve_all(engine, arrivals, gw.RESTART, 10, 10, pool=200_000, arena=64 << 20, page_log2=12)
    assert tot["pages_resident"] <= npages // 4
    for _, qid, _, _ in arrivals:
        assert np.array_equal(dev[qid][0], host[qid][0]) and np.array_equal(dev[qid][1], host[qid][1]), qid
