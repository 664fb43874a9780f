"""Walks placed per source (gw_walks_init_sources) are started by the
persistent kernel from a per-source table -- walk i is source entry i / wps,
id i, hop 0 -- instead of from walk records.  The same walks loaded as records
(gw_walks_load) take the record path.  Both must give the oracle's per-source
counts bit for bit, whatever the number of walks per source is against the
walks a wavefront claims (a claim can span many sources or sit inside one).
"""
import numpy as np
import pytest

import graphwalker_amd as gw
import oracle as O

pytestmark = pytest.mark.gpu


def oprog(p: gw.Program) -> O.Program:
    return O.Program(p.kind, p.L, p.stop_thr, p.nvertices, p.stream, p.seed)


@pytest.fixture(scope="module")
def graphs(engine):
    return O.rmat_graph(12, 1 << 16, seed=7), O.rmat_graph(12, 1 << 16, seed=8)


def expected(prog, off, col, sources, field, wps):
    """(field, vertex, count) triples sorted by (field, vertex), from the oracle."""
    s_, v_, c_ = [], [], []
    for k, (s, f) in enumerate(zip(sources, field)):
        w = O.make_walks(np.full(wps, s, np.uint32), sources_=np.full(wps, f, np.uint32),
                         ids=k * wps + np.arange(wps))
        c = O.replay(oprog(prog), off, col, w)["counts"]
        nz = np.nonzero(c)[0]
        s_.append(np.full(nz.shape[0], f, np.uint32))
        v_.append(nz.astype(np.uint32))
        c_.append(c[nz])
    s, v, c = np.concatenate(s_), np.concatenate(v_), np.concatenate(c_)
    # a field named by several sources (fields = vertices, a repeated source) sums
    key = (s.astype(np.uint64) << 32) | v
    uk, inv = np.unique(key, return_inverse=True)
    cs = np.zeros(uk.shape[0], np.uint64)
    np.add.at(cs, inv, c)
    return (uk >> 32).astype(np.uint32), (uk & 0xFFFFFFFF).astype(np.uint32), cs.astype(np.uint32)


def program(kind, n, chunk):
    prog = gw.program(kind, 20, 0.15, n, seed=3, recorder=gw.REC_SOURCE_COUNT | gw.REC_NO_FINAL,
                      schedule="persistent", layout="ext", source_method=gw.SRC_PARTITIONED)
    prog.walk_chunk = chunk
    return prog


@pytest.mark.parametrize("nsrc,wps", [(1, 5000), (7, 3), (50, 64), (33, 700), (300, 17), (64, 1)])
@pytest.mark.parametrize("kind", [gw.STOP, gw.RESTART, gw.JUMP])
@pytest.mark.parametrize("chunk", [64, 512])
def test_generated_starts_equal_records_and_oracle(engine, graphs, nsrc, wps, kind, chunk):
    (off, col), _ = graphs
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    sources = O.sources(5, n, nsrc)
    # STOP: the field is the source index (msppr's layout); RESTART: the vertex,
    # which is also the restart target; JUMP: the vertex, fields left implicit
    field = np.arange(nsrc, dtype=np.uint32) if kind == gw.STOP else sources
    prog = program(kind, n, chunk)
    engine.walks_init_sources(sources, wps, source_field=None if kind == gw.JUMP else field)
    st = engine.run(prog)
    got = engine.source_counts()
    engine.walks_load(O.make_walks(np.repeat(sources, wps), sources_=np.repeat(field, wps)))
    st2 = engine.run(prog)
    ref = engine.source_counts()
    assert st["steps"] == st2["steps"] and st["visits"] == st2["visits"]
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    for a, b in zip(got, expected(prog, off, col, sources, field, wps)):
        assert np.array_equal(a, b)
    assert int(got[2].sum()) == st["visits"]


def test_generated_starts_follow_a_new_graph(engine, graphs):
    """The start table holds each source's row: a graph set after the walks
    were placed (same vertex count, other edges) must be the one walked."""
    (off, col), (off2, col2) = graphs
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    sources = O.sources(9, n, 40)
    field = np.arange(40, dtype=np.uint32)
    prog = program(gw.STOP, n, 512)
    engine.walks_init_sources(sources, 200, source_field=field)
    engine.run(prog)
    first = engine.source_counts()
    engine.graph_set(off2, col2)
    engine.run(prog)
    got = engine.source_counts()
    for a, b in zip(got, expected(prog, off2, col2, sources, field, 200)):
        assert np.array_equal(a, b)
    assert not all(np.array_equal(a, b) for a, b in zip(first, got))


def test_bucketing_ends_generated_starts(engine, graphs):
    """gw_bucket reorders the walk array: the walks are no longer a function
    of their position, so runs read the records again (and stay exact)."""
    (off, col), _ = graphs
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.graph_blocks([0, 100, 1000, n])
    sources = O.sources(11, n, 30)
    field = np.arange(30, dtype=np.uint32)
    prog = program(gw.STOP, n, 128)
    engine.walks_init_sources(sources, 150, source_field=field)
    engine.bucket(3)
    engine.run(prog)
    got = engine.source_counts()
    for a, b in zip(got, expected(prog, off, col, sources, field, 150)):
        assert np.array_equal(a, b)


def test_generated_starts_under_graph_replay(engine, graphs):
    """A captured run of generated walks replays to the eager result.  After
    a new graph the replay either is refused (the graph's buffers moved) or,
    with the buffers at the same addresses, walks the new graph -- its start
    table is rebuilt before the launch -- and equals the eager run."""
    (off, col), (off2, col2) = graphs
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    sources = O.sources(13, n, 25)
    field = np.arange(25, dtype=np.uint32)
    prog = program(gw.STOP, n, 256)
    engine.walks_init_sources(sources, 333, source_field=field)
    engine.run(prog)
    eager = engine.source_counts()
    engine.run_capture(prog)
    for _ in range(3):
        engine.run_replay()
        got = engine.source_counts()
        for a, b in zip(got, eager):
            assert np.array_equal(a, b)
        engine.run(prog)
    engine.graph_set(off2, col2)
    try:
        engine.run_replay()
    except gw.GwError as e:
        assert "changed since the capture" in str(e)
        return
    got = engine.source_counts()
    for a, b in zip(got, expected(prog, off2, col2, sources, field, 333)):
        assert np.array_equal(a, b)
