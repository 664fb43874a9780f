"""hipGraph replay (gw_run_capture / gw_run_replay) and the launch-shape
autotuner (gw_autotune): replays give the eager run's results, replay after
replay with eager work in between; the tuned program gives the same results."""
import numpy as np
import pytest

import graphwalker_amd as gw
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g12():
    return O.rmat_graph(12, 1 << 16, seed=7)


def _state(engine, rec):
    out = {}
    if rec & gw.REC_VERTEX_COUNT:
        out["counts"] = engine.counts().copy()
    if rec & gw.REC_WALK_SUMMARY:
        h, v = engine.summary()
        out["hash"], out["visits"] = h.copy(), v.copy()
    if rec & gw.REC_SOURCE_COUNT:
        out["pairs"] = engine.source_counts()
    return out


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        if k == "pairs":
            assert all(np.array_equal(x, y) for x, y in zip(a[k], b[k]))
        else:
            assert np.array_equal(a[k], b[k]), k


CASES = [
    ("persistent", "ext", gw.JUMP, gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY, gw.SRC_SORT),
    ("plain", "csr", gw.STOP, gw.REC_VERTEX_COUNT, gw.SRC_SORT),
    ("persistent", "csr", gw.STOP, gw.REC_SOURCE_COUNT, gw.SRC_PARTITIONED),
    ("persistent", "ext", gw.RESTART, gw.REC_SOURCE_COUNT | gw.REC_WALK_SUMMARY, gw.SRC_HASH),
]


@pytest.mark.parametrize("sched,layout,kind,rec,method", CASES)
def test_replay_equals_eager(engine, g12, sched, layout, kind, rec, method):
    off, col = g12
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    fields = np.arange(40, dtype=np.uint32)
    engine.walks_init_sources(O.sources(3, n, 40), 500, source_field=fields)
    prog = gw.program(kind, 12, 0.15, n, seed=31, recorder=rec, schedule=sched, layout=layout, source_method=method)
    eager_st = engine.run(prog)
    eager = _state(engine, rec)
    engine.run_capture(prog)
    for _ in range(3):
        st = engine.run_replay()
        assert st["steps"] == eager_st["steps"] and st["visits"] == eager_st["visits"]
        _same(_state(engine, rec), eager)
        # eager work of the program's own between replays, then the graph again
        engine.graph_read()
    # new walks of the same count and source fields: the graph runs them
    engine.walks_init_sources(O.sources(4, n, 40), 500, source_field=fields)
    st = engine.run_replay()
    got = _state(engine, rec)
    eager_st = engine.run(prog)
    assert st["steps"] == eager_st["steps"]
    _same(got, _state(engine, rec))
    # a different walk count changes what the graph launches with
    engine.walks_init_sources(O.sources(4, n, 40), 400, source_field=fields)
    with pytest.raises(gw.GwError, match="changed since the capture"):
        engine.run_replay()


def test_capture_refuses_host_decided_runs(engine, g12):
    off, col = g12
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.graph_blocks([0, 1000, 2000, n])
    engine.walks_init_sources(O.sources(3, n, 10), 100)
    with pytest.raises(gw.GwError, match="blocked"):
        engine.run_capture(gw.program(gw.STOP, 10, 0.15, n, seed=1, schedule="blocked"))
    with pytest.raises(gw.GwError, match="global sort-reduce"):
        engine.run_capture(gw.program(gw.STOP, 10, 0.15, n, seed=1, recorder=gw.REC_SOURCE_COUNT,
                                      source_method=gw.SRC_GLOBAL))


def test_autotune_keeps_results(engine, g12):
    off, col = g12
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.walks_init_per_vertex(2)
    prog = gw.program(gw.JUMP, 10, 0.15, n, seed=8, recorder=gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY,
                      schedule="persistent", layout="ext")
    engine.run(prog)
    ref = _state(engine, prog.recorder)
    best, ms = engine.autotune(prog, reps=1)
    assert ms > 0 and best.block_threads in (64, 128, 256) and best.waves_per_simd in (2, 4, 6, 8, 12, 16)
    assert best.walk_chunk in (0, 128, 512, 2048, 8192)
    engine.run(best)
    _same(_state(engine, prog.recorder), ref)


def test_autotune_picks_block_size_for_blocked_runs(engine):
    # the blocked schedule's block table is tuned too and left installed; the
    # results are those of the eager run with any table
    off, col = O.rmat_graph(16, 1 << 20, seed=3)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.graph_blocks([0, 30000, n])
    engine.walks_init_sources(O.sources(5, n, 64), 200)
    prog = gw.program(gw.RESTART, 10, 0.15, n, seed=12, recorder=gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY,
                      schedule="blocked", layout="ext")
    engine.run(prog)
    ref = _state(engine, prog.recorder)
    best, ms = engine.autotune(prog, reps=1)
    nb = engine.graph_info()["nblocks"]
    assert ms > 0 and nb in (16, 4, 1)  # 2^12, 2^14 or 2^16 vertices per block
    engine.run(best)
    _same(_state(engine, prog.recorder), ref)
