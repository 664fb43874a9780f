"""Host walk pools (gw_walks_load_host): the records stay in host memory and a
run steps them window by window.  Walk ids key the random streams, so every
result must equal the oracle's replay and the same records loaded into HBM,
bit for bit, whatever the window size."""
import numpy as np
import pytest

import graphwalker_amd as gw
import oracle as O

pytestmark = pytest.mark.gpu

VARIANTS = [("plain", "csr"), ("persistent", "csr"), ("persistent", "ext"), ("blocked", "ext")]


def oprog(p: gw.Program) -> O.Program:
    return O.Program(p.kind, p.L, p.stop_thr, p.nvertices, p.stream, p.seed)


@pytest.fixture(scope="module")
def g12():
    return O.rmat_graph(12, 1 << 16, seed=7)


def _walks(n, nv, kind, seed):
    rng = np.random.default_rng(seed)
    v = rng.integers(0, nv, n).astype(np.uint32)
    hops = (np.arange(n) % 3) * 11 if kind == gw.RESTART_JOINT else np.zeros(n, np.uint32)
    return O.make_walks(v, sources_=v, hops=hops)  # ids 0..n-1 in order


@pytest.mark.parametrize("window", [3000, 10_000, 1 << 20])
@pytest.mark.parametrize("sched,layout", VARIANTS)
@pytest.mark.parametrize("kind", [gw.STOP, gw.RESTART_JOINT, gw.JUMP])
def test_pool_equals_oracle(engine, g12, kind, sched, layout, window):
    off, col = g12
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.graph_blocks([0, 700, 1500, 2600, n])
    walks = _walks(10_000, n, kind, seed=kind + 11)
    engine.walks_load_host(walks, window)
    prog = gw.program(kind, 11, 0.15, n, seed=0xC0FFEE, recorder=gw.REC_VERTEX_COUNT | gw.REC_WALK_SUMMARY,
                      schedule=sched, layout=layout)
    st = engine.run(prog)
    ref = O.replay(oprog(prog), off, col, walks, counts=True, per_walk=True)
    assert st["pool_windows"] == -(-10_000 // min(window, 10_000))
    assert st["steps"] == ref["steps"] and st["visits"] == ref["visits"]
    assert np.array_equal(engine.counts(), ref["counts"])
    assert np.array_equal(engine.walks_final(), ref["final"])
    h, vis = engine.summary()
    assert np.array_equal(h, ref["hash"]) and np.array_equal(vis, ref["walk_visits"])
    # a second run of the pool repeats the first (recorders cleared once per run, not per window)
    st2 = engine.run(prog)
    assert st2["steps"] == st["steps"] and np.array_equal(engine.counts(), ref["counts"])


def test_pool_events_equal_device_walks(engine, g12):
    off, col = g12
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    walks = _walks(5000, n, gw.STOP, seed=3)
    cap = 9
    prog = gw.program(gw.STOP, 12, 0.15, n, seed=77, recorder=gw.REC_EVENTS, event_cap=cap, schedule="persistent")
    engine.walks_load(walks)
    engine.run(prog)
    ev_dev, (h_dev, v_dev) = engine.events(cap), engine.summary()
    engine.walks_load_host(walks, 1234)
    engine.run(prog)
    h, v = engine.summary()
    assert np.array_equal(h, h_dev) and np.array_equal(v, v_dev)
    # entries beyond a walk's visit count are never written: compare the written ones
    written = np.arange(cap)[None, :] < np.minimum(v_dev, cap)[:, None]
    assert written.sum() > 0 and np.array_equal(engine.events(cap)[written], ev_dev[written])


def test_pool_source_counts_hash(engine, g12):
    off, col = g12
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    rng = np.random.default_rng(4)
    v = rng.integers(0, n, 8000).astype(np.uint32)
    walks = O.make_walks(v, sources_=(np.arange(8000) // 400).astype(np.uint32))
    prog = gw.program(gw.STOP, 15, 0.15, n, seed=5, recorder=gw.REC_SOURCE_COUNT, schedule="persistent", layout="ext")
    engine.walks_load(walks)
    engine.run(prog)
    ref = engine.source_counts()
    engine.walks_load_host(walks, 2500)
    st = engine.run(prog)
    assert st["source_method"] == gw.SRC_HASH and st["pool_windows"] == 4
    got = engine.source_counts()
    assert all(np.array_equal(a, b) for a, b in zip(got, ref))
    with pytest.raises(gw.GwError, match="GW_SRC_HASH"):
        engine.run(gw.program(gw.STOP, 15, 0.15, n, seed=5, recorder=gw.REC_SOURCE_COUNT,
                              source_method=gw.SRC_PARTITIONED))


def test_pool_pinned_in_place_and_host_graph(engine, g12):
    off, col = g12
    n = off.shape[0] - 1
    walks = _walks(6000, n, gw.RESTART, seed=8)
    pool = engine.host_empty(walks.shape, gw.WALK_DTYPE)
    pool[:] = walks
    engine.graph_set_host(off, col, 64 << 10)  # out-of-core graph and host pool together
    engine.walks_load_host(pool, 2048, keep_final=False)
    prog = gw.program(gw.RESTART, 10, 0.15, n, seed=21, recorder=gw.REC_VERTEX_COUNT)
    st = engine.run(prog)
    ref = O.replay(oprog(prog), off, col, walks, counts=True)
    assert st["schedule"] == gw.SCHED_BLOCKED and st["blocks_loaded"] > 0 and st["pool_windows"] == 3
    assert st["steps"] == ref["steps"] and np.array_equal(engine.counts(), ref["counts"])
    with pytest.raises(gw.GwError, match="keeps no final records"):
        engine.walks_final()


def test_pool_invalid(engine, g12):
    off, col = g12
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    walks = _walks(100, n, gw.STOP, seed=1)
    bad = walks.copy()
    bad["id"][5] = 6
    with pytest.raises(gw.GwError, match="id != its position"):
        engine.walks_load_host(bad, 10)
    engine.walks_load_host(walks, 10)
    prog = gw.program(gw.STOP, 5, 0.15, n, seed=1, schedule="persistent")
    with pytest.raises(gw.GwError, match="window by window"):
        engine.run_capture(prog)
    with pytest.raises(gw.GwError, match="host pool"):
        engine.bucket(1)
    # device walks replace the pool
    engine.walks_load(walks)
    engine.run(prog)
    assert engine.walks_final().shape[0] == 100


def test_walks_outside_a_new_graph_are_refused(engine, g12):
    off, col = g12
    engine.graph_set(off, col)
    engine.walks_load(_walks(100, off.shape[0] - 1, gw.STOP, seed=2))
    soff, scol = O.rmat_graph(6, 1 << 9, seed=1)  # 64 vertices: most walks start outside
    engine.graph_set(soff, scol)
    with pytest.raises(gw.GwError, match="outside the graph"):
        engine.run(gw.program(gw.STOP, 5, 0.15, 64, seed=1))
