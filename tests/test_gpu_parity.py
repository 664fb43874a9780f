"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact for everything (graphs, trajectories, counts, step totals): the
walk programs are integer-only and keyed by (seed, walk id, hop).
"""
import numpy as np
import pytest

import graphwalker_amd as gw
import oracle as O
from oracle import fp64

pytestmark = pytest.mark.gpu

KINDS = [gw.JUMP, gw.STOP, gw.PROB, gw.RESTART_JOINT, gw.RESTART]
SCHEDS = [gw.SCHED_PLAIN, gw.SCHED_PERSISTENT, gw.SCHED_BLOCKED]
# (schedule, layout) pairs the engine runs
VARIANTS = [("plain", "csr"), ("persistent", "csr"), ("persistent", "ext"), ("blocked", "csr"), ("blocked", "ext")]


def oprog(p: gw.Program) -> O.Program:
    return O.Program(p.kind, p.L, p.stop_thr, p.nvertices, p.stream, p.seed)


@pytest.fixture(scope="module")
def small_graph(engine):
    off, col = O.rmat_graph(12, 1 << 16, seed=7)
    return off, col


@pytest.mark.parametrize("scale,m,seed", [(4, 40, 3), (10, 9000, 5), (16, 1 << 20, 1), (20, 1 << 24, 2)])
def test_rmat_device_graph_equals_oracle(engine, scale, m, seed):
    engine.graph_rmat(scale, m, seed)
    off, col = engine.graph_read()
    roff, rcol = O.rmat_graph(scale, m, seed)
    assert np.array_equal(off, roff)
    assert np.array_equal(col, rcol)


def _walks(n, nv, kind, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.integers(0, nv, n).astype(np.uint32)
    hops = (np.arange(n) % 3) * 11 if kind == gw.RESTART_JOINT else np.zeros(n, np.uint32)
    w = O.make_walks(v, sources_=v, hops=hops)
    return w


@pytest.mark.parametrize("variant", VARIANTS, ids=["-".join(v) for v in VARIANTS])
@pytest.mark.parametrize("kind", KINDS)
def test_step_programs_bit_exact(engine, small_graph, kind, variant):
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    if variant[0] == "blocked":
        engine.graph_blocks([0, 100, 700, 1024, 2500, 4095, n])
    walks = _walks(20_000, n, kind, seed=kind)
    engine.walks_load(walks)
    prog = gw.program(kind, 11, 0.15, n, seed=0xDEADBEEF12345, recorder=gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY,
                      schedule=variant[0], layout=variant[1])
    st = engine.run(prog)
    assert st["layout"] == gw.LAYOUT_NAMES[variant[1]]
    if variant[0] == "blocked":
        assert st["rounds"] > 1
    ref = O.replay(oprog(prog), off, col, walks, per_walk=True)
    assert st["steps"] == ref["steps"] and st["visits"] == ref["visits"]
    assert np.array_equal(engine.counts(), ref["counts"])
    h, v = engine.summary()
    assert np.array_equal(v, ref["walk_visits"])
    assert np.array_equal(h, ref["hash"])
    fin = engine.walks_final()
    assert np.array_equal(fin, ref["final"])


def test_config1_restart_ppr(engine):
    """R-MAT 2^16 / 2^20, seed 1; 100k restart walks from vertex 0, p = 0.15,
    10 hops: 1,000,000 steps exactly, counts bit-exact, L1 to FP64 within the
    Monte-Carlo bound."""
    engine.graph_rmat(16, 1 << 20, 1)
    engine.graph_blocks(np.arange(0, (1 << 16) + 1, 4096))
    off, col = O.rmat_graph(16, 1 << 20, 1)
    W, L = 100_000, 10
    engine.walks_init_sources([0], W)
    prog = gw.program(gw.RESTART, L, 0.15, 1 << 16, seed=1, recorder=gw.REC_VERTEX_COUNT)
    for sched, layout in VARIANTS:
        prog.schedule = gw.SCHED_NAMES[sched]
        prog.layout = gw.LAYOUT_NAMES[layout]
        st = engine.run(prog)
        assert st["steps"] == W * L and st["visits"] == W * L
        c = engine.counts()
        ref = O.replay(oprog(prog), off, col, O.make_walks(np.zeros(W, np.uint32)))
        assert np.array_equal(c, ref["counts"])
    exp = fp64.expected_visits(O.RESTART, off, col, 0, W, L, 0.15)
    l1 = fp64.l1_normalised(c, exp)
    assert l1 < fp64.mc_l1_bound(exp, W, L, L)
    print(f"[config1] L1 = {l1:.4f}, bound {fp64.mc_l1_bound(exp, W, L, L):.4f}")


def test_config1_l1_below_001_with_enough_walks(engine):
    """The Monte-Carlo estimate converges to the FP64 finite-horizon vector:
    on the config-1 graph, 10 M restart walks bring the normalised L1 below
    0.01 (100 k walks give ~0.09, inside their own MC bound)."""
    engine.graph_rmat(16, 1 << 20, 1)
    off, col = engine.graph_read()
    W, L = 10_000_000, 10
    engine.walks_init_sources([0], W)
    st = engine.run(gw.program(gw.RESTART, L, 0.15, 1 << 16, seed=11, recorder=gw.REC_VERTEX_COUNT))
    assert st["steps"] == W * L
    exp = fp64.expected_visits(O.RESTART, off, col, 0, W, L, 0.15)
    l1 = fp64.l1_normalised(engine.counts(), exp)
    print(f"[config1-10M] L1 = {l1:.5f}")
    assert l1 < 0.01


def _per_source(s, v, c, nsrc, n):
    dense = np.zeros((nsrc, n), dtype=np.uint64)
    dense[s, v] = c
    return dense


def test_batched_ppr_per_source_fp64(engine):
    """The batched multi-source PPR path (stop 0.15, cap 20, 2000 walks per
    source, exact per-(source, vertex) counts by the partitioned method): each
    source's normalised vector against its own FP64 expectation, within the
    Monte-Carlo bound of 2000 walks, on the single-source PPR graph."""
    engine.graph_rmat(16, 1 << 20, 1)
    off, col = engine.graph_read()
    n = off.shape[0] - 1
    S, W, L = 128, 2000, 20
    src = O.sources(31, n, S)
    engine.walks_init_sources(src, W, source_field=np.arange(S, dtype=np.uint32))
    prog = gw.program(gw.STOP, L, 0.15, n, seed=17, recorder=gw.REC_SOURCE_COUNT | gw.REC_NO_FINAL,
                      source_method=gw.SRC_PARTITIONED)
    st = engine.run(prog)
    assert st["source_method"] == gw.SRC_PARTITIONED
    dense = _per_source(*engine.source_counts(), S, n)
    assert np.all(dense.sum(axis=1) > 0) and st["visits"] == int(dense.sum())
    expect = fp64.expected_visits_sources(O.STOP, off, col, src, W, L, 0.15)
    l1s, bounds = [], []
    for k in range(S):
        exp = expect[k]
        l1s.append(fp64.l1_normalised(dense[k], exp))
        bounds.append(fp64.mc_l1_bound(exp, W, exp.sum() / W, L))
    l1s, bounds = np.array(l1s), np.array(bounds)
    print(f"[ppr-batched] mean L1 {l1s.mean():.4f}, max L1/bound {np.max(l1s / bounds):.3f}")
    assert np.all(l1s < bounds)


def test_batched_ppr_l1_below_001_with_enough_walks(engine):
    """Two sources that spread (several out-edges) need about 10^7 walks each
    for a per-source L1 below 0.01 at cap 20; 2000 walks give 0.3-0.6."""
    engine.graph_rmat(16, 1 << 20, 1)
    off, col = engine.graph_read()
    n = off.shape[0] - 1
    deg = np.diff(off.astype(np.int64))
    cand = O.sources(3, n, 64)
    src = cand[deg[cand] >= 4][:2]
    assert src.shape[0] == 2
    W, L = 20_000_000, 20
    engine.walks_init_sources(src, W, source_field=np.arange(2, dtype=np.uint32))
    st = engine.run(gw.program(gw.STOP, L, 0.15, n, seed=23, recorder=gw.REC_SOURCE_COUNT | gw.REC_NO_FINAL))
    dense = _per_source(*engine.source_counts(), 2, n)
    assert st["visits"] == int(dense.sum())
    for k in range(2):
        exp = fp64.expected_visits(O.STOP, off, col, int(src[k]), W, L, 0.15)
        l1 = fp64.l1_normalised(dense[k], exp)
        print(f"[ppr-20M] source {src[k]} (out-degree {deg[src[k]]}): L1 = {l1:.5f}")
        assert l1 < 0.01


@pytest.mark.parametrize("tb,wps,chunk", [(64, 1, 64), (128, 3, 0), (256, 8, 8192), (192, 5, 2048)])
def test_launch_shape_independence(engine, small_graph, tb, wps, chunk):
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    walks = _walks(7_777, n, gw.STOP, seed=9)
    engine.walks_load(walks)
    prog = gw.program(gw.STOP, 20, 0.15, n, seed=3, recorder=gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY,
                      schedule=gw.SCHED_PERSISTENT, block_threads=tb, waves_per_simd=wps, walk_chunk=chunk)
    engine.run(prog)
    c1, (h1, _) = engine.counts(), engine.summary()
    prog.schedule = gw.SCHED_PLAIN
    engine.run(prog)
    assert np.array_equal(engine.counts(), c1)
    assert np.array_equal(engine.summary()[0], h1)
    # the blocked schedule claims walks the same way
    engine.graph_blocks([0, 700, 2500, n])
    prog.schedule = gw.SCHED_BLOCKED
    engine.run(prog)
    assert np.array_equal(engine.counts(), c1)
    assert np.array_equal(engine.summary()[0], h1)


@pytest.mark.parametrize("chunk", [32, 100, 8256])
def test_walk_chunk_invalid(engine, small_graph, chunk):
    off, col = small_graph
    engine.graph_set(off, col)
    engine.walks_load(_walks(100, off.shape[0] - 1, gw.STOP, seed=1))
    with pytest.raises(gw.GwError, match="walk_chunk"):
        engine.run(gw.program(gw.STOP, 5, 0.15, off.shape[0] - 1, seed=1, walk_chunk=chunk))


def test_edge_cases(engine):
    # dangling-only graph, empty walk set, walks already at the cap
    off = np.zeros(9, dtype=np.uint64)
    col = np.zeros(0, dtype=np.uint32)
    engine.graph_set(off, col)
    engine.walks_load(np.zeros(0, dtype=gw.WALK_DTYPE))
    st = engine.run(gw.program(gw.STOP, 5, 0.15, 8, seed=1))
    assert st["steps"] == 0 and st["visits"] == 0
    w = O.make_walks(np.arange(8), hops=[0, 5, 9, 0, 0, 0, 0, 0])
    engine.walks_load(w)
    for kind in (gw.JUMP, gw.STOP, gw.RESTART):
        prog = gw.program(kind, 5, 0.15, 8, seed=2, recorder=gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY)
        st = engine.run(prog)
        ref = O.replay(oprog(prog), off, col, w, per_walk=True)
        assert (st["steps"], st["visits"]) == (ref["steps"], ref["visits"])
        assert np.array_equal(engine.counts(), ref["counts"])
        assert np.array_equal(engine.walks_final(), ref["final"])
    # STOP on a dangling-only graph: each live walk visits once and dies
    assert ref is not None


def test_invalid_inputs_raise(engine, small_graph):
    off, col = small_graph
    engine.graph_set(off, col)
    bad = O.make_walks([1 << 20])
    with pytest.raises(gw.GwError):
        engine.walks_load(bad)
    with pytest.raises(gw.GwError):
        engine.run(gw.program(gw.PROB, 1, 0.0, 10, seed=1))
    with pytest.raises(gw.GwError):
        engine.graph_blocks([0, 5, 5, off.shape[0] - 1])
    engine.walks_load(O.make_walks([1, 2], sources_=[1, 1 << 30]))
    with pytest.raises(gw.GwError):
        engine.run(gw.program(gw.RESTART, 4, 0.15, 10, seed=1))


@pytest.mark.parametrize("nb", [1, 5, 64, 700])
def test_bucketing_is_a_stable_counting_sort(engine, small_graph, nb):
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    rng = np.random.default_rng(nb)
    cuts = np.unique(rng.integers(1, n, nb - 1)) if nb > 1 else np.array([], dtype=np.int64)
    bf = np.concatenate([[0], cuts, [n]]).astype(np.uint32)
    engine.graph_blocks(bf)
    for nw in (0, 1, 4095, 4097, 100_003):
        w = O.make_walks(rng.integers(0, n, nw).astype(np.uint32), ids=np.arange(nw))
        if nw:
            engine.walks_load(w)
        else:
            engine.walks_load(np.zeros(0, dtype=gw.WALK_DTYPE))
        cnt = engine.bucket(bf.shape[0] - 1)
        ref, rcnt = O.bucket_stable(w, bf)
        assert np.array_equal(cnt, rcnt[:-1])
        assert np.array_equal(engine.walks_current(), ref)


@pytest.mark.parametrize("variant", [("plain", "csr"), ("persistent", "ext"), ("blocked", "ext")])
@pytest.mark.parametrize("kind", [gw.STOP, gw.RESTART_JOINT, gw.PROB])
def test_event_log_is_the_trajectory(engine, small_graph, kind, variant):
    """The event log replays updateInfo exactly: walk w's k-th visit."""
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    if variant[0] == "blocked":
        engine.graph_blocks([0, 512, 1024, 3000, n])
    walks = _walks(3000, n, kind, seed=11)
    engine.walks_load(walks)
    cap = 12
    prog = gw.program(kind, 11, 0.15, n, seed=77, recorder=gw.REC_EVENTS, schedule=variant[0], layout=variant[1],
                      event_cap=cap)
    engine.run(prog)
    ev = engine.events(cap)
    _, vis = engine.summary()
    op = oprog(prog)
    for i in range(0, 3000, 7):
        tv, th, _ = O.trace(op, off, col, walks[i])
        k = min(len(tv), cap)
        assert vis[walks[i]["id"]] == len(tv)
        row = ev[walks[i]["id"], :k]
        assert np.array_equal(row & 0xFFFFFFFF, tv[:k]) and np.array_equal(row >> 32, th[:k])


@pytest.mark.parametrize("variant", [("plain", "csr"), ("persistent", "ext"), ("blocked", "ext")])
@pytest.mark.parametrize("kind", [gw.STOP, gw.RESTART, gw.JUMP, gw.PROB])
@pytest.mark.parametrize("cap", [1, 3])
def test_event_log_keeps_last_visit(engine, small_graph, kind, variant, cap):
    """REC_EVENTS_LAST: the first cap - 1 visits, then the latest visit in the
    last slot -- with cap 1 each walk's last (vertex, hop), the visit an
    updateInfo that looks only at hop L - 1 needs (graphlet.cpp:46-52)."""
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    if variant[0] == "blocked":
        engine.graph_blocks([0, 512, 1024, 3000, n])
    walks = _walks(3000, n, kind, seed=12)
    engine.walks_load(walks)
    prog = gw.program(kind, 4, 0.15, n, seed=78, recorder=gw.REC_EVENTS | gw.REC_EVENTS_LAST, schedule=variant[0],
                      layout=variant[1], event_cap=cap)
    engine.run(prog)
    ev = engine.events(cap)
    _, vis = engine.summary()
    op = oprog(prog)
    for i in range(0, 3000, 5):
        tv, th, _ = O.trace(op, off, col, walks[i])
        wid = walks[i]["id"]
        assert vis[wid] == len(tv)
        if len(tv) == 0:
            continue
        k = min(len(tv), cap)
        exp_v = np.concatenate([tv[:k - 1], tv[-1:]])
        exp_h = np.concatenate([th[:k - 1], th[-1:]])
        row = ev[wid, :k]
        assert np.array_equal(row & 0xFFFFFFFF, exp_v) and np.array_equal(row >> 32, exp_h), (i, kind, cap)


def test_events_last_needs_events(engine, small_graph):
    off, col = small_graph
    engine.graph_set(off, col)
    engine.walks_load(_walks(10, off.shape[0] - 1, gw.STOP, seed=1))
    with pytest.raises(gw.GwError, match="GW_REC_EVENTS_LAST needs GW_REC_EVENTS"):
        engine.run(gw.program(gw.STOP, 4, 0.15, recorder=gw.REC_EVENTS_LAST))


def _expected_source_counts(prog, off, col, sources, field, wps):
    out_s, out_v, out_c = [], [], []
    for k, (s, f) in enumerate(zip(sources, field)):
        w = O.make_walks(np.full(wps, s, np.uint32), sources_=np.full(wps, f, np.uint32),
                         ids=k * wps + np.arange(wps))
        c = O.replay(oprog(prog), off, col, w)["counts"]
        nz = np.nonzero(c)[0]
        out_s.append(np.full(nz.shape[0], f, np.uint32))
        out_v.append(nz.astype(np.uint32))
        out_c.append(c[nz])
    s, v, c = np.concatenate(out_s), np.concatenate(out_v), np.concatenate(out_c)
    order = np.lexsort((v, s))
    return s[order], v[order], c[order]


SRC_METHODS = [gw.SRC_SORT, gw.SRC_HASH, gw.SRC_GLOBAL, gw.SRC_PARTITIONED]


@pytest.mark.parametrize("method", SRC_METHODS, ids=["sort", "hash", "global", "partitioned"])
@pytest.mark.parametrize("variant", [("plain", "csr"), ("persistent", "ext"), ("persistent", "csr"), ("blocked", "ext")])
@pytest.mark.parametrize("kind", [gw.STOP, gw.RESTART])
def test_source_counts_bit_exact(engine, small_graph, kind, variant, method):
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    if variant[0] == "blocked":
        engine.graph_blocks([0, 64, 1000, 2048, n])
    sources = O.sources(3, n, 50)
    # STOP: the source field is the source index (msppr.cpp:59); RESTART: the vertex
    field = np.arange(50, dtype=np.uint32) if kind == gw.STOP else sources
    wps = 300
    engine.walks_init_sources(sources, wps, source_field=field)
    prog = gw.program(kind, 20, 0.15, n, seed=3, recorder=gw.REC_SOURCE_COUNT | gw.REC_WALK_SUMMARY,
                      schedule=variant[0], layout=variant[1], source_method=method)
    st = engine.run(prog)
    s, v, c = engine.source_counts()
    es, ev, ec = _expected_source_counts(prog, off, col, sources, field, wps)
    assert np.array_equal(s, es) and np.array_equal(v, ev) and np.array_equal(c, ec)
    assert int(c.sum()) == st["visits"]
    used = {gw.SRC_SORT: gw.SRC_PARTITIONED}.get(method, method)
    assert st["source_method"] == used


def test_source_counts_partitioned_flush_is_exact(engine):
    # JUMP teleports uniformly, so each source's 1500 x 40 visits cover far more
    # distinct vertices than one LDS table holds: the sources flush partial
    # pairs and are summed by the split sort-reduce.
    off, col = O.rmat_graph(16, 1 << 19, seed=4)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    sources = O.sources(9, n, 3)
    field = np.array([0, 7, 2], np.uint32)
    wps = 1500
    engine.walks_init_sources(sources, wps, source_field=field)
    prog = gw.program(gw.JUMP, 40, 0.15, n, seed=5, recorder=gw.REC_SOURCE_COUNT, source_method=gw.SRC_PARTITIONED)
    st = engine.run(prog)
    assert st["source_method"] == gw.SRC_PARTITIONED and st["flushed_sources"] == 3
    s, v, c = engine.source_counts()
    es, ev, ec = _expected_source_counts(prog, off, col, sources, field, wps)
    assert np.array_equal(s, es) and np.array_equal(v, ev) and np.array_equal(c, ec)
    # a second run on the same context reuses the regions and gives the same pairs
    st2 = engine.run(prog)
    s2, v2, c2 = engine.source_counts()
    assert st2["steps"] == st["steps"]
    assert np.array_equal(s2, es) and np.array_equal(v2, ev) and np.array_equal(c2, ec)


@pytest.mark.parametrize("schedule,layout", [("persistent", "ext"), ("plain", "csr")])
def test_source_counts_partitioned_large_regions_exact(engine, schedule, layout):
    # Regions with more keys than an aggregation workgroup holds in registers
    # (12288) but few distinct vertices are counted tile by tile without a
    # flush; sources on sinks have no region keys and are finished before the
    # workgroups start; the rest take the register path.
    off, col = O.rmat_graph(10, 1 << 14, seed=6)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    sinks = np.nonzero(np.diff(off) == 0)[0][:3].astype(np.uint32)
    sources = np.r_[O.sources(2, n, 5), sinks].astype(np.uint32)
    field = np.arange(sources.shape[0], dtype=np.uint32)
    wps = 8000
    engine.walks_init_sources(sources, wps, source_field=field)
    prog = gw.program(gw.STOP, 20, 0.15, n, seed=7, recorder=gw.REC_SOURCE_COUNT | gw.REC_NO_FINAL,
                      schedule=schedule, layout=layout, source_method=gw.SRC_PARTITIONED)
    st = engine.run(prog)
    assert st["source_method"] == gw.SRC_PARTITIONED and st["flushed_sources"] == 0
    s, v, c = engine.source_counts()
    es, ev, ec = _expected_source_counts(prog, off, col, sources, field, wps)
    assert np.array_equal(s, es) and np.array_equal(v, ev) and np.array_equal(c, ec)
    region = np.bincount(es, weights=np.where(ev == sources[es], 0, ec), minlength=sources.shape[0])
    assert region.max() > 12288 and sinks.shape[0] > 0 and (region[5:] == 0).all()


def test_source_counts_partitioned_all_sources_without_keys(engine):
    # every source is a vertex with no out-edges: no source has region keys,
    # the classify kernel writes each one's single pair and the aggregation
    # workgroups find an empty work list
    off, col = O.rmat_graph(10, 1 << 14, seed=6)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    sources = np.nonzero(np.diff(off) == 0)[0][:40].astype(np.uint32)
    field = np.arange(sources.shape[0], dtype=np.uint32)
    engine.walks_init_sources(sources, 50, source_field=field)
    prog = gw.program(gw.STOP, 20, 0.15, n, seed=7, recorder=gw.REC_SOURCE_COUNT | gw.REC_NO_FINAL,
                      schedule="persistent", layout="ext", source_method=gw.SRC_PARTITIONED)
    st = engine.run(prog)
    s_, v_, c_ = engine.source_counts()
    assert np.array_equal(s_, field) and np.array_equal(v_, sources) and (c_ == 50).all()
    assert st["visits"] == 50 * sources.shape[0] and st["flushed_sources"] == 0


def test_source_counts_region_overflow_falls_back(engine):
    # PROB walks are uncapped; regions are sized for 64 visits per walk.  Source
    # field 0 starts on a sink-free cycle (about 128 visits per walk at p = 1/128),
    # field 1 on a sink (one visit per walk): field 0's region overflows while the
    # whole log does not.  GW_SRC_SORT falls back to the global sort-reduce;
    # GW_SRC_PARTITIONED reports the overflow.
    n = 200
    deg = np.r_[np.ones(100, np.uint64), np.zeros(100, np.uint64)]
    off = np.r_[np.zeros(1, np.uint64), np.cumsum(deg)].astype(np.uint64)
    col = ((np.arange(100) + 1) % 100).astype(np.uint32)
    engine.graph_set(off, col)
    w = O.make_walks(np.r_[np.zeros(100, np.uint32), np.full(10_000, 150, np.uint32)],
                     sources_=np.r_[np.zeros(100, np.uint32), np.ones(10_000, np.uint32)])
    engine.walks_load(w)
    prog = gw.program(gw.PROB, 0, 1.0 / 128, n, seed=21, recorder=gw.REC_SOURCE_COUNT, source_method=gw.SRC_SORT)
    st = engine.run(prog)
    assert st["source_method"] == gw.SRC_GLOBAL
    assert st["visits"] > 100 * 64
    got = engine.source_counts()
    ref = O.replay(oprog(prog), off, col, w[:100])["counts"]
    nz = np.nonzero(ref)[0]
    assert np.array_equal(got[0][:nz.shape[0]], np.zeros(nz.shape[0], np.uint32))
    assert np.array_equal(got[1][:nz.shape[0]], nz) and np.array_equal(got[2][:nz.shape[0]], ref[nz])
    assert got[0].shape[0] == nz.shape[0] + 1 and got[1][-1] == 150 and got[2][-1] == 10_000
    with pytest.raises(gw.GwError, match="outgrew"):
        engine.run(gw.program(gw.PROB, 0, 1.0 / 128, n, seed=21, recorder=gw.REC_SOURCE_COUNT,
                              source_method=gw.SRC_PARTITIONED))


def test_source_table_overflow_is_an_error(engine, small_graph):
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.walks_init_sources(O.sources(3, n, 20), 200, source_field=np.arange(20, dtype=np.uint32))
    with pytest.raises(gw.GwError, match="table full"):
        engine.run(gw.program(gw.STOP, 20, 0.15, n, seed=3, recorder=gw.REC_SOURCE_COUNT, table_log2=6,
                              source_method=gw.SRC_HASH))


@pytest.mark.parametrize("method", [gw.SRC_HASH, gw.SRC_SORT])
def test_source_counts_wide_keys(engine, method):
    # fields >= 2^31 with 13-bit vertex ids: 45-bit keys, past what a packed
    # hash entry holds, so the table switches to two-word entries
    off, col = O.rmat_graph(13, 1 << 17, seed=8)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    sources = O.sources(4, n, 6)
    field = (np.uint32(0x80000000) + np.arange(6, dtype=np.uint32) * np.uint32(0x1000000)).astype(np.uint32)
    wps = 400
    engine.walks_init_sources(sources, wps, source_field=field)
    prog = gw.program(gw.STOP, 20, 0.15, n, seed=12, recorder=gw.REC_SOURCE_COUNT, source_method=method)
    st = engine.run(prog)
    s, v, c = engine.source_counts()
    es, ev, ec = _expected_source_counts(prog, off, col, sources, field, wps)
    assert np.array_equal(s, es) and np.array_equal(v, ev) and np.array_equal(c, ec)
    assert int(c.sum()) == st["visits"]


def _topk(s, v, c, k):
    order = np.lexsort((v, -c.astype(np.int64), s))  # source asc, count desc, vertex asc
    s, v, c = s[order], v[order], c[order]
    rank = np.arange(s.shape[0]) - np.searchsorted(s, s, side="left")
    keep = rank < k
    return s[keep], v[keep], c[keep]


@pytest.mark.parametrize("method", SRC_METHODS, ids=["sort", "hash", "global", "partitioned"])
@pytest.mark.parametrize("k", [1, 5, 1024])
def test_source_topk(engine, small_graph, method, k):
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    sources = O.sources(3, n, 30)
    field = np.arange(30, dtype=np.uint32)
    wps = 300
    engine.walks_init_sources(sources, wps, source_field=field)
    prog = gw.program(gw.STOP, 20, 0.15, n, seed=3, recorder=gw.REC_SOURCE_COUNT, source_method=method)
    engine.run(prog)
    s, v, c = engine.source_topk(k)
    es, ev, ec = _topk(*_expected_source_counts(prog, off, col, sources, field, wps), k)
    assert np.array_equal(s, es) and np.array_equal(v, ev) and np.array_equal(c, ec)


def _np_topk(counts, k):
    order = np.lexsort((np.arange(counts.shape[0]), -counts.astype(np.int64)))[:k]
    return order.astype(np.uint32), counts[order]


@pytest.mark.parametrize("L,k", [(10, 1), (10, 100), (10, 5000), (3, 3000), (3, 1 << 12), (3, 1 << 20)])
def test_counts_topk_equals_numpy(engine, L, k):
    # L = 3 leaves most counts small: many ties at the threshold, and k beyond
    # the visited vertices (threshold 0); k above n returns all n vertices
    off, col = O.rmat_graph(12, 1 << 15, seed=5)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.walks_init_per_vertex(3)
    engine.run(gw.program(gw.STOP, L, 0.15, n, seed=2, recorder=gw.REC_VERTEX_COUNT))
    counts = engine.counts()
    v, c = engine.counts_topk(k)
    ev, ec = _np_topk(counts, min(k, n))
    assert np.array_equal(v, ev) and np.array_equal(c, ec)


def test_counts_topk_large_counts(engine):
    # counts spread over all four digits of the radix select (walks pile up on hubs)
    off, col = O.rmat_graph(16, 1 << 20, seed=9)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.walks_init_per_vertex(64)
    engine.run(gw.program(gw.JUMP, 40, 0.15, n, seed=4, recorder=gw.REC_VERTEX_COUNT, schedule="persistent"))
    counts = engine.counts()
    assert counts.max() >= 1 << 16
    for k in (1, 777, 1 << 14, n):
        v, c = engine.counts_topk(k)
        ev, ec = _np_topk(counts, k)
        assert np.array_equal(v, ev) and np.array_equal(c, ec), k


def test_counts_topk_invalid(engine):
    off, col = O.rmat_graph(8, 1 << 10, seed=1)
    engine.graph_set(off, col)
    with pytest.raises(gw.GwError, match="bad arguments"):
        engine.counts_topk(0)


@pytest.mark.parametrize("sched,layout", VARIANTS)
def test_no_final_keeps_recorders(engine, small_graph, sched, layout):
    # GW_REC_NO_FINAL drops only the final records: counts and summaries are unchanged
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.walks_init_sources(O.sources(2, n, 30), 300)
    rec = gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY
    prog = gw.program(gw.STOP, 12, 0.15, n, seed=4, recorder=rec, schedule=sched, layout=layout)
    st = engine.run(prog)
    counts, (h, v), fin = engine.counts().copy(), engine.summary(), engine.walks_final()
    prog_nf = gw.program(gw.STOP, 12, 0.15, n, seed=4, recorder=rec | gw.REC_NO_FINAL, schedule=sched, layout=layout)
    st2 = engine.run(prog_nf)
    assert st2["steps"] == st["steps"] and np.array_equal(engine.counts(), counts)
    h2, v2 = engine.summary()
    assert np.array_equal(h2, h) and np.array_equal(v2, v)
    with pytest.raises(gw.GwError, match="no final records"):
        engine.walks_final()
    engine.run(prog)
    assert np.array_equal(engine.walks_final(), fin)


def _csr_from_edges(n, edges):
    src = np.array([e[0] for e in edges], dtype=np.int64)
    dst = np.array([e[1] for e in edges], dtype=np.uint32)
    order = np.lexsort((dst, src))
    off = np.zeros(n + 1, dtype=np.uint64)
    off[1:] = np.cumsum(np.bincount(src, minlength=n))
    return off, dst[order].astype(np.uint32)


UNUSUAL = {
    "one-vertex-self-loop": (1, [(0, 0)]),
    "star-1001": (1001, [(0, v) for v in range(1, 1001)] + [(v, 0) for v in range(1, 1001)]),
    "dups-and-loops-37": (37, [(v, (v * 7) % 37) for v in range(37)] * 3 + [(v, v) for v in range(0, 37, 4)]
                          + [(5, 6)] * 20),
}


@pytest.mark.parametrize("variant", VARIANTS, ids=["-".join(v) for v in VARIANTS])
@pytest.mark.parametrize("name", sorted(UNUSUAL))
def test_unusual_graphs_all_kinds(engine, name, variant):
    # a single vertex with a self-loop, a high-degree star with an odd vertex
    # count, duplicate edges and self-loops: every kind equals the oracle
    n, edges = UNUSUAL[name]
    off, col = _csr_from_edges(n, edges)
    engine.graph_set(off, col)
    if n > 2:
        engine.graph_blocks([0, n // 3, n])
    v = np.repeat(np.arange(n, dtype=np.uint32), 3)
    for kind in KINDS:
        hops = (np.arange(v.shape[0]) % 3) * 7 if kind == gw.RESTART_JOINT else np.zeros(v.shape[0], np.uint32)
        walks = O.make_walks(v, sources_=v, hops=hops)
        engine.walks_load(walks)
        prog = gw.program(kind, 7, 0.15, n, seed=kind + 40, recorder=gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY,
                          schedule=variant[0], layout=variant[1])
        st = engine.run(prog)
        ref = O.replay(oprog(prog), off, col, walks, counts=True, per_walk=True)
        assert (st["steps"], st["visits"]) == (ref["steps"], ref["visits"]), (name, kind)
        assert np.array_equal(engine.counts(), ref["counts"]), (name, kind)
        assert np.array_equal(engine.walks_final(), ref["final"]), (name, kind)
        h, vis = engine.summary()
        assert np.array_equal(h, ref["hash"]) and np.array_equal(vis, ref["walk_visits"]), (name, kind)


def _pairs_of(walks, off, col, prog):
    """Exact (field, vertex, count) of arbitrary loaded walks: the oracle's
    counts per field, sorted by field then vertex."""
    out = []
    for f in np.unique(walks["source"]):
        c = O.replay(oprog(prog), off, col, walks[walks["source"] == f])["counts"]
        nz = np.nonzero(c)[0]
        out.append((np.full(nz.shape[0], f, np.uint32), nz.astype(np.uint32), c[nz]))
    return tuple(np.concatenate([o[i] for o in out]) for i in range(3))


@pytest.mark.parametrize("variant", [("plain", "csr"), ("persistent", "ext"), ("blocked", "ext")])
@pytest.mark.parametrize("kind", [gw.STOP, gw.RESTART, gw.RESTART_JOINT])
def test_source_counts_mixed_starts(engine, small_graph, kind, variant):
    # fields whose walks start at different vertices and hops: each field's
    # designated vertex (its first walk's start) is counted apart by the
    # direct-region path; every method gives the oracle's pairs
    off, col = small_graph
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    if variant[0] == "blocked":
        engine.graph_blocks([0, 512, 1024, 3000, n])
    rng = np.random.default_rng(kind + 40)
    nw = 6000
    field = rng.integers(0, 37, nw).astype(np.uint32)
    start = np.where(rng.random(nw) < 0.5, (field * 97) % n, rng.integers(0, n, nw)).astype(np.uint32)
    hops = (rng.integers(0, 3, nw) * (11 if kind == gw.RESTART_JOINT else 2)).astype(np.uint32)
    walks = O.make_walks(start, sources_=field, hops=hops)
    engine.walks_load(walks)
    for method in (gw.SRC_PARTITIONED, gw.SRC_GLOBAL, gw.SRC_HASH):
        prog = gw.program(kind, 11, 0.15, n, seed=91, recorder=gw.REC_SOURCE_COUNT, schedule=variant[0],
                          layout=variant[1], source_method=method)
        st = engine.run(prog)
        assert st["source_method"] == method
        es, ev, ec = _pairs_of(walks, off, col, prog)
        s, v, c = engine.source_counts()
        assert np.array_equal(s, es) and np.array_equal(v, ev) and np.array_equal(c, ec), method
