"""SimRank on the device: gw_graph_transpose, gw_walks_init_pairs and
gw_meeting_counts against the oracle's trajectories and a brute-force
first-meeting count (apps/simrank.cpp:71-84)."""
import numpy as np
import pytest

import graphwalker_amd as gw
import oracle as O

pytestmark = pytest.mark.gpu


def _transpose(off, col):
    n = off.shape[0] - 1
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    order = np.lexsort((src, col))
    toff = np.zeros(n + 1, dtype=np.uint64)
    toff[1:] = np.cumsum(np.bincount(col, minlength=n))
    return toff, src[order].astype(np.uint32)


@pytest.mark.parametrize("scale,m,seed", [(4, 40, 3), (12, 1 << 16, 7), (16, 1 << 20, 2)])
def test_transpose_equals_numpy(engine, scale, m, seed):
    off, col = O.rmat_graph(scale, m, seed=seed)
    engine.graph_set(off, col)
    engine.graph_transpose()
    toff, tcol = engine.graph_read()
    eoff, ecol = _transpose(off, col)
    assert np.array_equal(toff, eoff) and np.array_equal(tcol, ecol)
    engine.graph_transpose()  # twice = the original graph
    roff, rcol = engine.graph_read()
    assert np.array_equal(roff, off) and np.array_equal(rcol, col)


def _paths(prog, off, col, start, R, L, id0):
    """Path by offset of walks id0.. id0+R-1 from `start` (hop i*L)."""
    P = L - 1
    out = np.full((R, P), 0xFFFFFFFF, dtype=np.uint64)
    for i in range(R):
        w = O.make_walks([start], sources_=[start], hops=[i * L], ids=[id0 + i])[0]
        tv, th, _ = O.trace(prog, off, col, w)
        out[i, (th - i * L).astype(np.int64)] = tv
    return out


def _first_meetings(A, B, L):
    counts = np.zeros(L, dtype=np.uint64)
    eq = (A[:, None, :] == B[None, :, :]) & (A[:, None, :] != 0xFFFFFFFF)
    any_ = eq.any(axis=2)
    first = eq.argmax(axis=2)
    np.add.at(counts, first[any_], 1)
    return counts


@pytest.mark.parametrize("transposed,L", [(False, 11), (True, 11), (False, 2), (False, 5), (True, 32)])
def test_meeting_counts_equal_brute_force(engine, transposed, L):
    off, col = O.rmat_graph(10, 1 << 13, seed=4)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    if transposed:
        engine.graph_transpose()
        off, col = _transpose(off, col)
    R = 60
    pairs = O.sources(5, n, 6).reshape(3, 2)
    a, b = pairs[:, 0].copy(), pairs[:, 1].copy()
    a[2] = b[2]  # a pair of one vertex meets at offset 0 every time
    engine.walks_init_pairs(a, b, R, L)
    prog = gw.program(gw.RESTART_JOINT, L, 0.15, n, seed=6, recorder=gw.REC_EVENTS, event_cap=L - 1)
    st = engine.run(prog)
    assert st["steps"] == 3 * 2 * R * (L - 1)
    got = engine.meeting_counts(3, R, L)
    op = O.Program(prog.kind, prog.L, prog.stop_thr, prog.nvertices, prog.stream, prog.seed)
    for k in range(3):
        A = _paths(op, off, col, a[k], R, L, 2 * k * R)
        B = _paths(op, off, col, b[k], R, L, 2 * k * R + R)
        assert np.array_equal(got[k], _first_meetings(A, B, L)), k
    assert got[2, 0] == R * R


def test_meeting_counts_large_tiles(engine):
    # R large enough that the b-side paths take several LDS tiles
    off, col = O.rmat_graph(9, 1 << 12, seed=2)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    R, L = 1500, 11
    a, b = np.array([3], np.uint32), np.array([7], np.uint32)
    engine.walks_init_pairs(a, b, R, L)
    prog = gw.program(gw.RESTART_JOINT, L, 0.15, n, seed=9, recorder=gw.REC_EVENTS, event_cap=L - 1)
    engine.run(prog)
    got = engine.meeting_counts(1, R, L)
    op = O.Program(prog.kind, prog.L, prog.stop_thr, prog.nvertices, prog.stream, prog.seed)
    A = _paths(op, off, col, 3, R, L, 0)
    B = _paths(op, off, col, 7, R, L, R)
    assert np.array_equal(got[0], _first_meetings(A, B, L))


def test_meeting_counts_invalid(engine):
    off, col = O.rmat_graph(9, 1 << 12, seed=2)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    engine.walks_init_pairs(np.array([1], np.uint32), np.array([2], np.uint32), 10, 11)
    with pytest.raises(gw.GwError, match="npairs x 2R"):
        engine.meeting_counts(2, 10, 11)
    with pytest.raises(gw.GwError, match="at most 31"):
        engine.meeting_counts(1, 10, 40)
    with pytest.raises(gw.GwError, match="out of range"):
        engine.walks_init_pairs(np.array([n], np.uint32), np.array([2], np.uint32), 10, 11)


@pytest.mark.parametrize("R,same", [(1024, True), (1, False), (333, False)])
def test_meeting_counts_sparse_join_edges(engine, R, same):
    # the sparse join's limits: R at its maximum with every pair meeting at
    # offset 0 (a == b), a single walk per side, a hub-heavy small graph
    off, col = O.rmat_graph(7, 1 << 11, seed=5)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    L = 6
    a = np.array([3, 9], np.uint32)
    b = np.array([3 if same else 17, 40], np.uint32)
    engine.walks_init_pairs(a, b, R, L)
    prog = gw.program(gw.RESTART_JOINT, L, 0.15, n, seed=13, recorder=gw.REC_EVENTS, event_cap=L - 1)
    engine.run(prog)
    got = engine.meeting_counts(2, R, L)
    op = O.Program(prog.kind, prog.L, prog.stop_thr, prog.nvertices, prog.stream, prog.seed)
    for k in range(2):
        A = _paths(op, off, col, a[k], R, L, 2 * k * R)
        B = _paths(op, off, col, b[k], R, L, 2 * k * R + R)
        assert np.array_equal(got[k], _first_meetings(A, B, L)), k
    if same:
        assert got[0, 0] == R * R


@pytest.mark.parametrize("R,copies", [(512, 64), (2048, 8)])
def test_meeting_rates_match_fp64_chain(engine, R, copies):
    # distributional parity: first-meeting counts / R^2 against the exact
    # pair chain of two independent restart walks; R <= 1024 takes the sparse
    # join, R > 1024 the brute-force one.  Each copy of a pair is an
    # independent estimate; a copy's count sums R^2 indicators that share
    # walks, so its variance is at most f/R^2 + 2f/R.
    from oracle import fp64
    off, col = O.rmat_graph(6, 1 << 9, seed=2)
    n = off.shape[0] - 1
    engine.graph_set(off, col)
    L = 9
    pairs = [(3, 9), (0, 1), (5, 5)]
    a = np.repeat(np.array([p[0] for p in pairs], np.uint32), copies)
    b = np.repeat(np.array([p[1] for p in pairs], np.uint32), copies)
    engine.walks_init_pairs(a, b, R, L)
    prog = gw.program(gw.RESTART_JOINT, L, 0.15, n, seed=21, recorder=gw.REC_EVENTS, event_cap=L - 1)
    engine.run(prog)
    got = engine.meeting_counts(a.shape[0], R, L).astype(np.float64) / (R * R)
    p = prog.stop_thr / 2.0 ** 32  # the coin restarts when the draw falls below stop_thr
    for k, (x, y) in enumerate(pairs):
        f = fp64.first_meeting(off, col, x, y, p, L - 1)
        est = got[k * copies:(k + 1) * copies, :L - 1].mean(axis=0)
        tol = 5.0 * np.sqrt((f / (R * R) + 2.0 * f / R) / copies) + 1e-9
        assert np.all(np.abs(est - f) <= tol), (x, y, est, f, tol)
        assert np.all(got[k * copies:(k + 1) * copies, L - 1] == 0)
        if x == y:
            assert np.all(est[1:] == 0.0) and est[0] == 1.0
    # the estimate of SimRank itself, sum_t c^t f_t, within the summed bound
    c = 0.8  # simrank.cpp:77 weighs a first meeting at offset l by 0.8^l
    for k, (x, y) in enumerate(pairs[:2]):
        f = fp64.first_meeting(off, col, x, y, p, L - 1)
        w = c ** np.arange(L - 1)
        est = (got[k * copies:(k + 1) * copies, :L - 1] @ w).mean()
        tol = 5.0 * float(w @ np.sqrt((f / (R * R) + 2.0 * f / R) / copies)) + 1e-9
        print(f"[simrank fp64] pair ({x},{y}) R={R}: estimate {est:.5f}, exact {float(w @ f):.5f}, bound {tol:.5f}")
        assert abs(est - float(w @ f)) <= tol
