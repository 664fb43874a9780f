"""Pins the CPU oracle before anything is compared with it.

Philox: Random123 known-answer vectors and a second implementation (rocRAND's
header engine, vectors in tests/golden/philox_vectors.json).
R-MAT / CSR: a pure-Python restatement on small cases, quadrant frequencies,
numpy lexsort equality.
Step programs: closed forms (cycles, sinks, p = 0 / p = 1), walk-length laws,
neighbour uniformity, FP64 finite-horizon expectations within the Monte-Carlo
bound, whole-graph replay == block-engine replay, stable bucketing ==
numpy stable sort.
"""
import json
import os

import numpy as np
import pytest
from scipy import stats

import oracle as O
from oracle import fp64

HERE = os.path.dirname(os.path.abspath(__file__))

# ---------------------------------------------------------------- Philox


def test_philox_random123_kat():
    # Random123 kat_vectors, philox4x32 with 10 rounds.
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, out in kat:
        assert O.philox(ctr, key)[0].tolist() == out


def test_philox_matches_rocrand_vectors():
    d = json.load(open(os.path.join(HERE, "golden", "philox_vectors.json")))
    assert len(d["vectors"]) == 64
    for v in d["vectors"]:
        assert O.philox(v["ctr"], v["key"])[0].tolist() == v["out"]


def _py_philox(c, k):
    c = list(c)
    k = list(k)
    M = 0xFFFFFFFF
    for r in range(10):
        if r:
            k = [(k[0] + 0x9E3779B9) & M, (k[1] + 0xBB67AE85) & M]
        p0 = 0xD2511F53 * c[0]
        p1 = 0xCD9E8D57 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k[0]) & M, p1 & M, ((p0 >> 32) ^ c[3] ^ k[1]) & M, p0 & M]
    return c


def test_philox_uniformity():
    ctr = np.zeros((20000, 4), dtype=np.uint32)
    ctr[:, 0] = np.arange(20000)
    out = O.philox(ctr, [7, 0]).reshape(-1)
    hist = np.bincount(out >> 28, minlength=16)
    assert stats.chisquare(hist).pvalue > 1e-4


# ------------------------------------------------------------- R-MAT/CSR


def _py_rmat_edge(e, scale, seed, ta, tab, tabc):
    s = d = 0
    r = None
    for lvl in range(scale):
        if lvl % 8 == 0:
            r = _py_philox([e & 0xFFFFFFFF, e >> 32, 0x524D4154, lvl // 8], [seed & 0xFFFFFFFF, seed >> 32])
        w = r[(lvl % 8) // 2]
        u = (w >> 16) if lvl % 2 else (w & 0xFFFF)
        if u < ta:
            q = (0, 0)
        elif u < tab:
            q = (0, 1)
        elif u < tabc:
            q = (1, 0)
        else:
            q = (1, 1)
        s = (s << 1) | q[0]
        d = (d << 1) | q[1]
    return s, d


def test_rmat_matches_python_restatement():
    ta, tab, tabc = O.rmat_thresholds(0.57, 0.19, 0.19)
    assert (ta, tab, tabc) == (37356, 49807, 62259)
    for scale in (5, 9, 17):
        src, dst = O.rmat_edges(scale, 300, seed=11, e_begin=(1 << 33) if scale == 17 else 0)
        for k in range(300):
            e = k + ((1 << 33) if scale == 17 else 0)
            assert (int(src[k]), int(dst[k])) == _py_rmat_edge(e, scale, 11, ta, tab, tabc)


def test_rmat_quadrant_frequencies():
    # Top-level bit pair of each edge follows (a, b, c, d).
    src, dst = O.rmat_edges(16, 200_000, seed=1)
    q = (src >> 15) * 2 + (dst >> 15)
    hist = np.bincount(q, minlength=4)
    exp = np.array([0.57, 0.19, 0.19, 0.05]) * src.shape[0]
    assert stats.chisquare(hist, exp).pvalue > 1e-4
    # Level bits are independent draws: low-level pair has the same law.
    q0 = (src & 1) * 2 + (dst & 1)
    assert stats.chisquare(np.bincount(q0, minlength=4), exp).pvalue > 1e-4


def test_csr_build_matches_lexsort():
    src, dst = O.rmat_edges(12, 50_000, seed=5)
    off, col = O.csr_build(1 << 12, src, dst)
    order = np.lexsort((dst, src))
    assert np.array_equal(col, dst[order])
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(src, minlength=1 << 12))]).astype(np.uint64))


def test_csr_rejects_bad_vertices():
    with pytest.raises(ValueError):
        O.csr_build(4, np.array([0, 5], np.uint32), np.array([1, 1], np.uint32))


# -------------------------------------------------------- step programs


def cycle(n):
    off = np.arange(n + 1, dtype=np.uint64)
    col = ((np.arange(n) + 1) % n).astype(np.uint32)
    return off, col


def test_cycle_p0_is_deterministic():
    # Out-degree 1 and p = 0: v_k = (s + k) mod n whatever the random stream.
    off, col = cycle(7)
    for kind in (O.JUMP, O.STOP, O.RESTART):
        prog = O.program(kind, 10, 0.0, 7, seed=99)
        tv, th, fin = O.trace(prog, off, col, O.make_walks([3])[0])
        assert tv.tolist() == [(3 + k) % 7 for k in range(10)]
        assert th.tolist() == list(range(10))
        assert int(fin["vertex"]) == (3 + 10) % 7 and int(fin["hop"]) == 10


def test_joint_period_and_restart():
    off, col = cycle(5)
    # p = 1: every step restarts at the source; period L = 4 records 3 visits.
    prog = O.program(O.RESTART_JOINT, 4, 1.0, 5, seed=1)
    w = O.make_walks([2], sources_=[2], hops=[8])[0]
    tv, th, fin = O.trace(prog, off, col, w)
    assert tv.tolist() == [2, 2, 2] and th.tolist() == [8, 9, 10]
    assert int(fin["hop"]) == 11
    # p = 0 on the cycle: L - 1 moves then stop.
    prog = O.program(O.RESTART_JOINT, 11, 0.0, 5, seed=1)
    tv, th, fin = O.trace(prog, off, col, O.make_walks([0], hops=[0])[0])
    assert tv.tolist() == [k % 5 for k in range(10)] and int(fin["hop"]) == 10


def sink_graph():
    # 0 -> 1, 1 -> 2, 2 has no out-edges (dangling)
    off = np.array([0, 1, 2, 2], dtype=np.uint64)
    col = np.array([1, 2], dtype=np.uint32)
    return off, col


def test_dead_end_rules():
    off, col = sink_graph()
    w = O.make_walks([0], sources_=[0])[0]
    # STOP dies at the sink: visits 0, 1, 2 then ends without moving.
    tv, _, fin = O.trace(O.program(O.STOP, 10, 0.0, 3, seed=4), off, col, w)
    assert tv.tolist() == [0, 1, 2] and int(fin["vertex"]) == 2 and int(fin["hop"]) == 2
    # PROB: same, no cap.
    tv, _, _ = O.trace(O.program(O.PROB, 1, 0.0, 3, seed=4), off, col, w)
    assert tv.tolist() == [0, 1, 2]
    # RESTART returns to the source at the sink: 0 1 2 0 1 2 ...
    tv, _, _ = O.trace(O.program(O.RESTART, 9, 0.0, 3, seed=4), off, col, w)
    assert tv.tolist() == [0, 1, 2] * 3
    # JUMP teleports from the sink: exactly L visits, every 3rd step random.
    prog = O.program(O.JUMP, 30, 0.0, 3, seed=4)
    tv, _, fin = O.trace(prog, off, col, w)
    assert len(tv) == 30 and int(fin["hop"]) == 30
    for k in range(29):
        if tv[k] != 2:
            assert tv[k + 1] == tv[k] + 1


def test_p1_never_follows_edges():
    off, col = cycle(6)
    tv, _, _ = O.trace(O.program(O.RESTART, 8, 1.0, 6, seed=2), off, col, O.make_walks([4])[0])
    assert tv.tolist() == [4] * 8
    tv, _, fin = O.trace(O.program(O.STOP, 8, 1.0, 6, seed=2), off, col, O.make_walks([4])[0])
    assert tv.tolist() == [4] and int(fin["hop"]) == 0


def _py_trace(kind, L, thr, N, seed, off, col, w):
    """Second, independent implementation of the step loop (pure Python)."""
    v, hop, src, wid = int(w["vertex"]), int(w["hop"]), int(w["source"]), int(w["id"])
    out = []
    while True:
        if kind in (O.JUMP, O.STOP, O.RESTART) and hop >= L:
            break
        out.append((v, hop))
        r = _py_philox([wid, hop, 0x57414C4B, 0], [seed & 0xFFFFFFFF, seed >> 32])
        b, deg = int(off[v]), int(off[v + 1] - off[v])
        if deg > 0 and r[0] >= thr:
            x = r[1] | (r[2] << 32)
            v = int(col[b + ((x * deg) >> 64)])
        elif kind == O.JUMP:
            v = (r[3] * N) >> 32
        elif kind in (O.STOP, O.PROB):
            break
        else:
            v = src
        hop += 1
        if kind == O.RESTART_JOINT and hop % L == L - 1:
            break
    return out


@pytest.mark.parametrize("kind", [O.JUMP, O.STOP, O.PROB, O.RESTART_JOINT, O.RESTART])
def test_trajectories_match_python_restatement(kind):
    off, col = O.rmat_graph(8, 2048, seed=21)
    N = 256
    prog = O.program(kind, 12, 0.15, N, seed=0x1234_5678_9ABC)
    rng = np.random.default_rng(kind)
    for i in range(40):
        v = int(rng.integers(0, N))
        w = O.make_walks([v], sources_=[v], hops=[0], ids=[int(rng.integers(0, 1 << 32))])[0]
        tv, th, _ = O.trace(prog, off, col, w)
        ref = _py_trace(kind, 12, prog.stop_thr, N, prog.seed, off, col, w)
        assert list(zip(tv.tolist(), th.tolist())) == ref


def test_jump_takes_exactly_L_steps():
    off, col = O.rmat_graph(10, 8000, seed=3)
    walks = O.make_walks(np.arange(1024))
    r = O.replay(O.program(O.JUMP, 10, 0.15, 1024, seed=8), off, col, walks, per_walk=True)
    assert r["steps"] == 10 * 1024 and r["visits"] == 10 * 1024
    assert np.all(r["final"]["hop"] == 10)
    assert int(r["counts"].sum()) == 10 * 1024


def test_stop_length_is_geometric():
    # On a graph without dead ends (cycle), STOP visits ~ min(Geom(0.15), L).
    off, col = cycle(101)
    L = 20
    walks = O.make_walks(np.zeros(200_000, dtype=np.uint32))
    r = O.replay(O.program(O.STOP, L, 0.15, 101, seed=5), off, col, walks, per_walk=True)
    k = r["walk_visits"].astype(np.int64)
    pk = np.array([0.85 ** (j - 1) * 0.15 for j in range(1, L)] + [0.85 ** (L - 1)])
    hist = np.bincount(k, minlength=L + 1)[1:]
    assert stats.chisquare(hist, pk * walks.shape[0]).pvalue > 1e-4
    # moves = final hop; a walk that died before the cap moved k - 1 times
    hop = r["final"]["hop"].astype(np.int64)
    assert r["steps"] == int(hop.sum()) and r["visits"] == int(k.sum())
    assert np.all(hop[k < L] == k[k < L] - 1)
    assert set(np.unique(hop[k == L]).tolist()) <= {L - 1, L}


def test_neighbour_choice_is_uniform_over_slots():
    # star: vertex 0 -> 1..40 (vertex 5 listed twice); others -> 0
    nb = list(range(1, 41)) + [5]
    off = np.array([0, len(nb)] + [len(nb) + i for i in range(1, 41)], dtype=np.uint64)
    col = np.array(nb + [0] * 40, dtype=np.uint32)
    walks = O.make_walks(np.zeros(100_000, dtype=np.uint32))
    r = O.replay(O.program(O.JUMP, 2, 0.0, 41, seed=17), off, col, walks)
    c = r["counts"][1:].astype(np.float64)
    exp = np.full(40, 1.0)
    exp[4] = 2.0
    exp = exp / exp.sum() * c.sum()
    assert stats.chisquare(c, exp).pvalue > 1e-4


def test_walk_ids_key_the_streams():
    off, col = O.rmat_graph(10, 8000, seed=3)
    prog = O.program(O.STOP, 10, 0.15, 1024, seed=8)
    a = O.make_walks(np.arange(64), ids=np.arange(64))
    b = a[::-1].copy()
    ra = O.replay(prog, off, col, a, per_walk=True)
    rb = O.replay(prog, off, col, b, per_walk=True)
    assert np.array_equal(ra["hash"], rb["hash"][::-1])
    assert np.array_equal(ra["counts"], rb["counts"])


# ------------------------------------------------ FP64 finite horizon


def test_restart_counts_match_fp64_within_mc_bound():
    off, col = O.rmat_graph(12, 1 << 16, seed=1)
    W, L, p = 200_000, 10, 0.15
    walks = O.make_walks(np.zeros(W, dtype=np.uint32))
    r = O.replay(O.program(O.RESTART, L, p, 1 << 12, seed=1), off, col, walks)
    exp = fp64.expected_visits(O.RESTART, off, col, 0, W, L, p)
    assert abs(exp.sum() - W * L) < 1e-6 * W * L
    l1 = fp64.l1_normalised(r["counts"], exp)
    bound = fp64.mc_l1_bound(exp, W, L, L)
    assert l1 < bound, (l1, bound)
    # Expected visits at the source in closed form for the first two hops:
    # 1 (hop 0) + P(back at 0 at hop 1) where hop-1 mass at 0 = p + (1-p) P[0,0].
    P, dang = fp64.transition(off, col)
    assert not dang[0]


def test_stop_and_jump_fp64():
    off, col = O.rmat_graph(11, 1 << 14, seed=2)
    n = 1 << 11
    W, L = 150_000, 8
    walks = O.make_walks(np.full(W, 3, dtype=np.uint32))
    for kind in (O.STOP, O.JUMP):
        r = O.replay(O.program(kind, L, 0.15, n, seed=9), off, col, walks)
        exp = fp64.expected_visits(kind, off, col, 3, W, L, 0.15)
        assert abs(r["visits"] - exp.sum()) < 5 * np.sqrt(exp.sum() * L)
        assert fp64.l1_normalised(r["counts"], exp) < fp64.mc_l1_bound(exp, W, exp.sum() / W, L)


def test_fp64_single_vertex_closed_form():
    # one vertex with a self-loop: STOP visits sum_{k<L} (1-p)^k per walk.
    off = np.array([0, 1], dtype=np.uint64)
    col = np.array([0], dtype=np.uint32)
    e = fp64.expected_visits(O.STOP, off, col, 0, 1.0, 6, 0.25)
    assert abs(e[0] - sum(0.75 ** k for k in range(6))) < 1e-12
    e = fp64.expected_visits(O.RESTART, off, col, 0, 1.0, 6, 0.25)
    assert abs(e[0] - 6.0) < 1e-12


# ------------------------------------------------ block engine, buckets


def test_bucket_stable_matches_numpy():
    rng = np.random.default_rng(0)
    w = O.make_walks(rng.integers(0, 1000, 5000))
    bf = np.array([0, 100, 350, 351, 900, 990], dtype=np.uint32)  # 990..999 out of range
    out, cnt = O.bucket_stable(w, bf)
    blk = np.searchsorted(bf, w["vertex"], side="right") - 1
    keep = w["vertex"] < bf[-1]
    order = np.argsort(blk[keep], kind="stable")
    assert np.array_equal(out, w[keep][order])
    assert cnt[-1] == (~keep).sum()


@pytest.mark.parametrize("kind", [O.JUMP, O.STOP, O.RESTART_JOINT, O.RESTART])
def test_block_engine_equals_whole_graph_replay(kind):
    off, col = O.rmat_graph(12, 1 << 15, seed=7)
    n = 1 << 12
    bf = np.array([0, 300, 1024, 1500, 2048, 4096], dtype=np.uint32)
    rng = np.random.default_rng(1)
    v = rng.integers(0, n, 3000).astype(np.uint32)
    walks = O.make_walks(v, hops=(np.arange(3000) % 5) * 11 if kind == O.RESTART_JOINT else None)
    prog = O.program(kind, 11, 0.15, n, seed=77)
    whole = O.replay(prog, off, col, walks)
    counts, steps, visits, rounds = O.block_engine(prog, off, col, bf, walks)
    assert np.array_equal(counts, whole["counts"])
    assert (steps, visits) == (whole["steps"], whole["visits"])
    assert rounds > 5


@pytest.mark.skipif(O.reference_app("rawrandomwalks") is None, reason="reference apps not built (oracle/build_ref.sh)")
def test_reference_engine_reads_written_graph(tmp_path):
    # the on-disk writer and the reference's reader agree: the reference's own
    # rawrandomwalks runs on the written graph, computes its block range and
    # reports its timers (the bench's reference CPU baseline takes this path)
    off, col = O.rmat_graph(10, 1 << 13, seed=3)
    base = str(tmp_path / "g")
    O.write_reference_graph(base, off, col)
    r = O.run_reference_rawrandomwalks(base, off.shape[0] - 1, 2000, 5, 2, 8, 4, timeout=60)
    assert r["00_runtime"] > 0 and r["5_exec_updates"] > 0
    blocks = np.loadtxt(base + "_GraphWalker/blocksize_8KB.blockrange", dtype=np.int64)
    assert blocks[0] == 0 and blocks[-1] == off.shape[0] - 1 and np.all(np.diff(blocks) > 0)


def test_fp64_first_meeting_closed_forms():
    # leaves 0 and 1 each point at hub 2, which loops on itself: with q = 1-p,
    # first meetings at offsets 0, 1, 2 are 0, q^2 and q^2 (1 - q^2); a == b
    # meets at offset 0 for sure
    off = np.array([0, 1, 2, 3], dtype=np.uint64)
    col = np.array([2, 2, 2], dtype=np.uint32)
    p = 0.15
    q = 1 - p
    f = fp64.first_meeting(off, col, 0, 1, p, 3)
    assert np.allclose(f, [0.0, q * q, q * q * (1 - q * q)], rtol=0, atol=1e-15)
    g = fp64.first_meeting(off, col, 1, 1, p, 4)
    assert g[0] == 1.0 and np.all(g[1:] == 0.0)
    # on any graph the first-meeting probabilities are a sub-distribution
    roff, rcol = O.rmat_graph(6, 1 << 9, seed=2)
    h = fp64.first_meeting(roff, rcol, 3, 9, p, 12)
    assert np.all(h >= 0) and h.sum() <= 1.0 + 1e-12


def test_fp64_first_meeting_matches_replayed_walks():
    # the bit-exact replay's first meetings (walk i of a side at hop i*L,
    # ids as gw_walks_init_pairs lays them out) agree with the chain within
    # the Monte-Carlo bound of copies independent R x R estimates
    roff, rcol = O.rmat_graph(6, 1 << 9, seed=2)
    n = roff.shape[0] - 1
    L, R, copies = 7, 100, 100
    prog = O.Program(O.RESTART_JOINT, L, int(0.15 * 2 ** 32), n, 0, 5)
    p = prog.stop_thr / 2.0 ** 32

    def paths(start, id0):
        out = np.full((R, L - 1), -1, dtype=np.int64)
        for i in range(R):
            w = O.make_walks([start], sources_=[start], hops=[i * L], ids=[id0 + i])[0]
            tv, th, _ = O.trace(prog, roff, rcol, w)
            out[i, (th - i * L).astype(np.int64)] = tv
        return out

    wrong = []
    for x, y in [(3, 9), (0, 1)]:
        f = fp64.first_meeting(roff, rcol, x, y, p, L - 1)
        est = np.zeros(L - 1)
        for c in range(copies):
            A, B = paths(x, 2 * c * R), paths(y, 2 * c * R + R)
            eq = A[:, None, :] == B[None, :, :]
            hit = eq.any(axis=2)
            np.add.at(est, eq.argmax(axis=2)[hit], 1.0)
        est /= copies * R * R
        tol = 5.0 * np.sqrt((f / (R * R) + 2.0 * f / R) / copies) + 1e-9
        assert np.all(np.abs(est - f) <= tol), (x, y, est, f, tol)
        assert f[1:].sum() > 0.05  # the pair meets often enough to test something
        wrong.append(np.any(np.abs(fp64.first_meeting(roff, rcol, x, y, 0.0, L - 1) - est) > tol))
    # the bound is tight enough that a chain without its restart coin fails
    assert any(wrong)


_REF_SKIP = pytest.mark.skipif(O.reference_app("visit_counts") is None or O.reference_app("simrank") is None,
                               reason="reference apps not built (oracle/build_ref.sh)")


@_REF_SKIP
@pytest.mark.parametrize("W", [100_000, 1_000_000])
def test_reference_engine_ppr_within_mc_bound(tmp_path, W):
    # both implementations against FP64 on the single-source PPR graph (R-MAT
    # 2^16 / 2^20, seed 1, source 0, stop 0.15, cap 10): the reference CPU
    # engine (its own RandomWalkwithStop and rand_r) and the Philox replay the
    # device matches bit for bit
    off, col = O.rmat_graph(16, 1 << 20, seed=1)
    n = off.shape[0] - 1
    base = str(tmp_path / "g")
    O.write_reference_graph(base, off, col)
    exp = fp64.expected_visits(O.STOP, off, col, 0, W, 10, 0.15)
    bound = fp64.mc_l1_bound(exp, W, exp.sum() / W, 10)
    ref = O.run_reference_visit_counts(base, n, 0, W, 10, 8, timeout=120)
    assert ref.shape == (n,) and ref[0] >= W  # every walk visits its source first
    l1_ref = fp64.l1_normalised(ref, exp)
    prog = O.Program(O.STOP, 10, int(0.15 * 2 ** 32), n, 0, 0x5EED)
    ours = O.replay(prog, off, col, O.make_walks(np.zeros(W, np.uint32)), counts=True)["counts"]
    l1_ours = fp64.l1_normalised(ours, exp)
    assert l1_ref < bound and l1_ours < bound, (l1_ref, l1_ours, bound)


@pytest.mark.skipif(O.reference_app("msppr_sources") is None, reason="reference apps not built (oracle/build_ref.sh)")
def test_reference_engine_multisource_ppr_within_mc_bound(tmp_path):
    # the CPU-baseline app for the batched PPR workload (the reference engine,
    # msppr's recorder, a list of uniform sources): it places and steps every
    # walk (visit total within the MC bound of the FP64 expectation over all
    # sources), and source 0's distribution -- exact below the recorder's
    # 1000-entry capacity on a 1024-vertex graph -- is within the MC bound of
    # the FP64 single-source vector
    off, col = O.rmat_graph(10, 1 << 14, seed=3)
    n = off.shape[0] - 1
    base = str(tmp_path / "g")
    O.write_reference_graph(base, off, col)
    src = O.sources(3, n, 16)
    W, L = 4000, 20
    ref = O.run_reference_msppr_sources(base, src, W, L, 8, timeout=120, blocksize_kb=1024)
    assert ref["00_runtime"] > 0
    exp_all = sum(fp64.expected_visits(O.STOP, off, col, int(s), W, L, 0.15).sum() for s in src)
    sd = np.sqrt(len(src) * W * L * L)  # a walk makes at most L visits
    assert abs(ref["visits"] - exp_all) <= 5 * sd / 4, (ref["visits"], exp_all)
    c = np.zeros(n, np.uint64)
    c[ref["source0_vertex"]] = ref["source0_count"]
    assert ref["source0_vertex"].shape[0] < 1000
    exp = fp64.expected_visits(O.STOP, off, col, int(src[0]), W, L, 0.15)
    bound = fp64.mc_l1_bound(exp, W, exp.sum() / W, L)
    assert fp64.l1_normalised(c, exp) < bound


@_REF_SKIP
def test_reference_engine_simrank_within_mc_bound(tmp_path):
    # the reference simrank app's score (first meetings weighed 0.8^l) against
    # the FP64 pair chain.  One pair's score averages R^2 walk-pair values in
    # [0, 1] that share walks: Var <= S/R^2 + 2S/R; over K pairs the mean error
    # has Var <= mean(S/R^2 + 2S/R) / K
    off, col = O.rmat_graph(6, 1 << 9, seed=2)
    base = str(tmp_path / "g")
    O.write_reference_graph(base, off, col)
    R, L = 1000, 11
    w = 0.8 ** np.arange(L - 1)
    rng = np.random.default_rng(4)
    pairs = [(int(a), int(b)) for a, b in rng.integers(0, 64, (12, 2)) if a != b]
    err, var = [], []
    for a, b in pairs:
        S = float(w @ fp64.first_meeting(off, col, a, b, 0.15, L - 1))
        s = O.run_reference_simrank(base, a, b, R, L, 8, timeout=60)
        v = S / (R * R) + 2.0 * S / R
        assert abs(s - S) <= 5.0 * np.sqrt(v) + 1e-6, (a, b, s, S)
        err.append(s - S)
        var.append(v)
    assert abs(np.mean(err)) <= 5.0 * np.sqrt(np.mean(var) / len(err))
    # a == b: every walk carries source a, so the app records all paths on the
    # a side (simrank.cpp:60-63) and reports 0, not the definition's 1 (the
    # device join counts R^2 first meetings at offset 0 for such a pair)
    assert O.run_reference_simrank(base, 5, 5, 100, L, 2, timeout=60) == 0.0


def test_fp64_sources_equals_single_source():
    # the batched FP64 expectation is the single-source one, source by source,
    # for every kind (duplicate sources and a dead end included)
    off, col = O.rmat_graph(8, 1 << 11, seed=5)
    n = off.shape[0] - 1
    deg = np.diff(off.astype(np.int64))
    src = np.array([0, 7, 7, int(np.argmin(deg)), n - 1])
    for kind in (O.JUMP, O.STOP, O.PROB, O.RESTART_JOINT, O.RESTART):
        many = fp64.expected_visits_sources(kind, off, col, src, 300.0, 9, 0.2)
        for k, s_ in enumerate(src):
            one = fp64.expected_visits(kind, off, col, int(s_), 300.0, 9, 0.2)
            assert np.allclose(many[k], one, rtol=1e-12, atol=1e-9), (kind, k)


def test_fp64_return_at_closed_forms():
    # a directed triangle 0 -> 1 -> 2 -> 0: every walk is back at hop 3 with
    # probability q^3, never at hop 2
    off = np.array([0, 1, 2, 3], dtype=np.uint64)
    col = np.array([1, 2, 0], dtype=np.uint32)
    q = 0.85
    assert np.allclose(fp64.return_at(off, col, 4, 0.15), q ** 3, rtol=0, atol=1e-15)
    assert np.allclose(fp64.return_at(off, col, 3, 0.15), 0.0, atol=1e-15)
    # brute force over every path of length 3 on a small random graph (CSR
    # slots with multiplicity, dead ends stop the walk)
    roff, rcol = O.rmat_graph(4, 60, seed=3)
    n = roff.shape[0] - 1
    deg = np.diff(roff.astype(np.int64))
    exp = np.zeros(n)
    for s in range(n):
        def walk(v, k, prob):
            if k == 3:
                if v == s:
                    exp[s] += prob
                return
            if deg[v] == 0:
                return
            for e in range(int(roff[v]), int(roff[v + 1])):
                walk(int(rcol[e]), k + 1, prob * q / deg[v])
        walk(s, 0, 1.0)
    assert np.allclose(fp64.return_at(roff, rcol, 4, 0.15), exp, rtol=1e-12, atol=1e-15)


def test_fp64_return_at_matches_replay():
    # the bit-exact replay's closed walks (a visit to the source at hop L-1)
    # within 5 sigma of the binomial bound, which a chain without the stop
    # coin would miss
    off, col = O.rmat_graph(6, 1 << 10, seed=4)
    n = off.shape[0] - 1
    L, W = 4, 40_000
    prog = O.Program(O.STOP, L, int(0.15 * 2 ** 32), n, 0, 9)
    p = prog.stop_thr / 2.0 ** 32
    starts = (np.arange(W) % n).astype(np.uint32)
    w = O.make_walks(starts, sources_=starts)
    closed = 0
    for i in range(W):
        tv, th, _ = O.trace(prog, off, col, w[i])
        closed += int(np.any((th == L - 1) & (tv == starts[i])))
    f = float(fp64.return_at(off, col, L, p)[starts].mean())
    tol = 5 * np.sqrt(f * (1 - f) / W)
    assert abs(closed / W - f) <= tol, (closed / W, f)
    assert abs(closed / W - float(fp64.return_at(off, col, L, 0.0)[starts].mean())) > tol


@pytest.mark.skipif(O.reference_app("graphlet") is None, reason="reference apps not built (oracle/build_ref.sh)")
def test_reference_engine_graphlet_within_mc_bound(tmp_path):
    # the reference CPU engine's closed-walk ratio (its own rand() sources and
    # rand_r coins) against the FP64 probability averaged over uniform sources
    off, col = O.rmat_graph(6, 1 << 10, seed=4)
    n = off.shape[0] - 1
    base = str(tmp_path / "g")
    O.write_reference_graph(base, off, col)
    R, L = 200_000, 4
    cnt = O.run_reference_graphlet(base, n, R, L, 8, timeout=120)
    f = float(fp64.return_at(off, col, L, 0.15).mean())
    assert abs(cnt / R - f) <= 5 * np.sqrt(f * (1 - f) / R), (cnt / R, f)


def test_fp64_stop_visits_pull_equals_sparse_chain():
    # the C pull over the transposed CSR against fp64.expected_visits_sources
    # (scipy), on a graph with dead ends, duplicate edges and self-loops
    off, col = O.rmat_graph(10, 1 << 13, seed=9)
    n = off.shape[0] - 1
    deg = np.diff(off.astype(np.int64))
    src_of = np.repeat(np.arange(n, dtype=np.uint32), deg)
    toff, tcol = O.csr_build(n, col, src_of)  # in-neighbour lists
    sources = np.array([0, 1, 5, 77, int(np.argmax(deg == 0)), 1023], dtype=np.uint32)
    got = O.fp64_stop_visits(off, toff, tcol, sources, 2000.0, 20, 0.15)
    exp = fp64.expected_visits_sources(O.STOP, off, col, sources, 2000.0, 20, 0.15)
    assert np.allclose(got, exp, rtol=1e-12, atol=1e-9)
    assert got[4, sources[4]] == 2000.0 and got[4].sum() == 2000.0  # a dead end: one visit per walk
