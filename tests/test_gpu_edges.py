"""The device CSR build from a host edge list (gw_graph_set_edges) against
the layouts it must reproduce: the reference converter's (rows in the
list's order, self-loops dropped, n = largest id + 1 -- the bytes
tests/test_frontend.py shows convert_to_csr writes for a list grouped by
source), a stable grouping by source for an unordered list, and the
oracle's CSR build (rows sorted) with sort_rows."""
import numpy as np
import pytest

import graphwalker_amd as gw
import oracle as O

pytestmark = pytest.mark.gpu


def _list(n, m, seed, grouped):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    src[: m // 20] = dst[: m // 20]  # self-loops
    src[m // 20: m // 10] = src[0]   # a heavy row with duplicates
    dst[m // 20: m // 10] = dst[1]
    if grouped:
        o = np.argsort(src, kind="stable")
        src, dst = src[o], dst[o]
    return src, dst


def test_converter_layout_from_a_list_grouped_by_source(engine):
    src, dst = _list(3000, 40_000, 1, grouped=True)
    engine.graph_set_edges(src, dst, drop_self_loops=True)
    off, col = engine.graph_read()
    keep = src != dst
    n = int(max(src[keep].max(), dst[keep].max())) + 1
    assert engine.n == n and engine.m == int(keep.sum())
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(src[keep], minlength=n))]).astype(np.uint64))
    assert np.array_equal(col, dst[keep])  # rows keep the list's order


@pytest.mark.parametrize("drop", [False, True])
def test_unordered_list_is_grouped_stably(engine, drop):
    src, dst = _list(5000, 200_000, 2, grouped=False)
    engine.graph_set_edges(src, dst, n=6000, drop_self_loops=drop)
    off, col = engine.graph_read()
    keep = (src != dst) if drop else np.ones(src.shape[0], bool)
    s, d = src[keep], dst[keep]
    o = np.argsort(s, kind="stable")
    assert engine.n == 6000
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(s, minlength=6000))]).astype(np.uint64))
    assert np.array_equal(col, d[o])


def test_sorted_rows_equal_the_oracle_csr(engine):
    src, dst = O.rmat_edges(14, 1 << 18, 5)
    engine.graph_set_edges(src, dst, n=1 << 14, sort_rows=True)
    off, col = engine.graph_read()
    roff, rcol = O.csr_build(1 << 14, src, dst)
    assert np.array_equal(off, roff) and np.array_equal(col, rcol)
    # and the device R-MAT build of the same edges
    engine.graph_rmat(14, 1 << 18, 5)
    doff, dcol = engine.graph_read()
    assert np.array_equal(off, doff) and np.array_equal(col, dcol)


def test_edge_list_refusals(engine):
    with pytest.raises(gw.GwError, match="not below n"):
        engine.graph_set_edges(np.array([0, 5], np.uint32), np.array([1, 2], np.uint32), n=4)
    engine.graph_set_edges(np.zeros(0, np.uint32), np.zeros(0, np.uint32), n=10)
    off, col = engine.graph_read()
    assert engine.n == 10 and col.shape == (0,) and np.all(off == 0)
