"""Out-of-core graphs (gw_graph_set_host): the CSR stays in host memory and
runs step it block by block through a capped device block cache.

Trajectories do not depend on which blocks are resident when (the random
stream is keyed by walk id and hop), so every result must equal the oracle's
replay on the whole graph, bit for bit, at any cache size.
"""
import numpy as np
import pytest

import graphwalker_amd as gw
import oracle as O

pytestmark = pytest.mark.gpu

KINDS = [gw.JUMP, gw.STOP, gw.PROB, gw.RESTART_JOINT, gw.RESTART]


def oprog(p: gw.Program) -> O.Program:
    return O.Program(p.kind, p.L, p.stop_thr, p.nvertices, p.stream, p.seed)


@pytest.fixture(scope="module")
def graph():
    return O.rmat_graph(12, 1 << 16, seed=7)  # 4096 vertices: 32 KiB offsets, 256 KiB columns


def _walks(n, nv, kind, seed):
    rng = np.random.default_rng(seed)
    v = rng.integers(0, nv, n).astype(np.uint32)
    hops = (np.arange(n) % 3) * 11 if kind == gw.RESTART_JOINT else np.zeros(n, np.uint32)
    return O.make_walks(v, sources_=v, hops=hops)


def _expect(prog, off, col, walks):
    return O.replay(oprog(prog), off, col, walks, counts=True, per_walk=True)


@pytest.mark.parametrize("cache_kib", [40, 96, 4096])
@pytest.mark.parametrize("kind", KINDS)
def test_streamed_runs_equal_oracle(engine, graph, kind, cache_kib):
    off, col = graph
    n = off.shape[0] - 1
    engine.graph_set_host(off, col, cache_kib << 10)
    walks = _walks(20_000, n, kind, seed=kind + 3)
    engine.walks_load(walks)
    prog = gw.program(kind, 11, 0.15, n, seed=0xBEEF, recorder=gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY)
    st = engine.run(prog)
    ref = _expect(prog, off, col, walks)
    assert st["schedule"] == gw.SCHED_BLOCKED and st["layout"] == gw.LAYOUT_CSR
    assert st["steps"] == ref["steps"] and st["visits"] == ref["visits"]
    assert np.array_equal(engine.counts(), ref["counts"])
    assert np.array_equal(engine.walks_final(), ref["final"])  # ids are 0..n-1 in load order
    h, vis = engine.summary()
    assert np.array_equal(h, ref["hash"]) and np.array_equal(vis, ref["walk_visits"])
    assert st["blocks_loaded"] >= 1 and st["h2d_bytes"] > 0 and st["cache_slots"] >= 1
    if cache_kib == 40:
        assert st["blocks_loaded"] > st["cache_slots"]  # the cache had to evict


def test_streamed_explicit_blocks_two_slots(engine, graph):
    off, col = graph
    n = off.shape[0] - 1
    bf = np.array([0, 100, 700, 1024, 2500, 3000, 3500, 4095, n], np.uint32)
    # room for exactly two slots of the largest block
    sizes = [(bf[k + 1] - bf[k] + 1) * 8 + (off[bf[k + 1]] - off[bf[k]]) * 4 for k in range(len(bf) - 1)]
    # a slot holds the largest block: offsets (256-byte aligned), then columns
    slot = max((((int(bf[k + 1] - bf[k]) + 1) * 8 + 255) // 256 * 256 + int(off[bf[k + 1]] - off[bf[k]]) * 4 + 255)
               // 256 * 256 for k in range(len(bf) - 1))
    engine.graph_set_host(off, col, 2 * slot + 100)
    engine.graph_blocks(bf)
    walks = _walks(10_000, n, gw.RESTART, seed=5)
    engine.walks_load(walks)
    prog = gw.program(gw.RESTART, 10, 0.15, n, seed=9, recorder=gw.REC_VERTEX_COUNT)
    st = engine.run(prog)
    ref = _expect(prog, off, col, walks)
    assert st["cache_slots"] == 2 and st["rounds"] > 1
    # walks start in every block, so every block is copied at least once
    assert st["blocks_loaded"] >= len(sizes) and st["h2d_bytes"] >= sum(int(x) for x in sizes)
    assert np.array_equal(engine.counts(), ref["counts"])


def test_streamed_events_and_source_counts(engine, graph):
    off, col = graph
    n = off.shape[0] - 1
    engine.graph_set_host(off, col, 64 << 10)
    walks = _walks(3000, n, gw.STOP, seed=11)
    engine.walks_load(walks)
    cap = 12
    prog = gw.program(gw.STOP, 11, 0.15, n, seed=77, recorder=gw.REC_EVENTS, event_cap=cap)
    engine.run(prog)
    ev = engine.events(cap)
    _, vis = engine.summary()
    op = oprog(prog)
    for i in range(0, 3000, 13):
        tv, th, _ = O.trace(op, off, col, walks[i])
        k = min(len(tv), cap)
        assert vis[walks[i]["id"]] == len(tv)
        row = ev[walks[i]["id"], :k]
        assert np.array_equal(row & 0xFFFFFFFF, tv[:k]) and np.array_equal(row >> 32, th[:k])
    # source counts: GW_SRC_SORT falls to the hash table on a host-resident graph
    sources = O.sources(3, n, 20)
    engine.walks_init_sources(sources, 200, source_field=np.arange(20, dtype=np.uint32))
    prog = gw.program(gw.STOP, 20, 0.15, n, seed=3, recorder=gw.REC_SOURCE_COUNT)
    st = engine.run(prog)
    assert st["source_method"] == gw.SRC_HASH
    s, v, c = engine.source_counts()
    out = []
    for k, src in enumerate(sources):
        w = O.make_walks(np.full(200, src, np.uint32), sources_=np.full(200, k, np.uint32), ids=k * 200 + np.arange(200))
        cnt = O.replay(oprog(prog), off, col, w)["counts"]
        nz = np.nonzero(cnt)[0]
        out.append((np.full(nz.shape[0], k, np.uint32), nz.astype(np.uint32), cnt[nz]))
    es, ev_, ec = (np.concatenate([o[i] for o in out]) for i in range(3))
    order = np.lexsort((ev_, es))
    assert np.array_equal(s, es[order]) and np.array_equal(v, ev_[order]) and np.array_equal(c, ec[order])


def test_streamed_pinned_arrays_in_place(engine, graph):
    off, col = graph
    n = off.shape[0] - 1
    hoff = engine.host_empty(off.shape, np.uint64)
    hcol = engine.host_empty(col.shape, np.uint32)
    hoff[:] = off
    hcol[:] = col
    engine.graph_set_host(hoff, hcol, 48 << 10)
    walks = _walks(5000, n, gw.JUMP, seed=2)
    engine.walks_load(walks)
    prog = gw.program(gw.JUMP, 10, 0.15, n, seed=4, recorder=gw.REC_VERTEX_COUNT)
    engine.run(prog)
    assert np.array_equal(engine.counts(), _expect(prog, off, col, walks)["counts"])
    roff, rcol = engine.graph_read()
    assert np.array_equal(roff, off) and np.array_equal(rcol, col)


def test_streamed_invalid_requests(engine, graph):
    off, col = graph
    n = off.shape[0] - 1
    engine.graph_set_host(off, col, 64 << 10)
    engine.walks_load(_walks(100, n, gw.STOP, seed=1))
    for kw in ({"schedule": "persistent"}, {"schedule": "plain"}, {"layout": "ext"}):
        with pytest.raises(gw.GwError, match="host-resident"):
            engine.run(gw.program(gw.STOP, 5, 0.15, n, seed=1, **kw))
    with pytest.raises(gw.GwError, match="GW_SRC_HASH"):
        engine.run(gw.program(gw.STOP, 5, 0.15, n, seed=1, recorder=gw.REC_SOURCE_COUNT,
                              source_method=gw.SRC_PARTITIONED))
    # (serving over a host-resident graph is supported: tests/test_gpu_serve_ooc.py)
    engine.graph_set_host(off, col, 1000)  # blocks cut to 250 bytes; the biggest vertex row does not fit
    engine.walks_load(_walks(100, n, gw.STOP, seed=1))
    with pytest.raises(gw.GwError, match="smaller than one block slot"):
        engine.run(gw.program(gw.STOP, 5, 0.15, n, seed=1))


@pytest.mark.parametrize("policy,prob", [(gw.BLOCKS_MIN_STEP, 1.0), (gw.BLOCKS_MIN_STEP, 0.2),
                                         (gw.BLOCKS_MAX_WEIGHT, 0.0)])
def test_block_policies_keep_results(engine, graph, policy, prob):
    # the load-pass block choice changes which blocks are copied when, never a trajectory
    off, col = graph
    n = off.shape[0] - 1
    engine.graph_set_host(off, col, 40 << 10)
    walks = _walks(20_000, n, gw.RESTART_JOINT, seed=6)
    engine.walks_load(walks)
    prog = gw.program(gw.RESTART_JOINT, 11, 0.15, n, seed=0xBEEF, recorder=gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY)
    ref = _expect(prog, off, col, walks)
    engine.stream_policy(policy, prob, seed=3)
    try:
        st = engine.run(prog)
        counts1 = engine.counts().copy()
        # the second run starts with the first one's blocks still cached: other loads, same results
        st2 = engine.run(prog)
    finally:
        engine.stream_policy(gw.BLOCKS_MAX_WALKS)
    assert st["steps"] == ref["steps"] and np.array_equal(counts1, ref["counts"])
    assert st2["steps"] == ref["steps"] and np.array_equal(engine.counts(), ref["counts"])
    h, vis = engine.summary()
    assert np.array_equal(h, ref["hash"]) and np.array_equal(vis, ref["walk_visits"])
    assert st["blocks_loaded"] > st["cache_slots"]


def test_block_policy_invalid(engine):
    with pytest.raises(gw.GwError, match="unknown policy"):
        engine.stream_policy(7)
    with pytest.raises(gw.GwError, match="unknown policy"):
        engine.stream_policy(gw.BLOCKS_MIN_STEP, 1.5)
