"""GPU runs through the C++ front end: an app on the GraphWalker API and
reference apps compiled unchanged against it (graphwalker_amd/bin/compat/, built
by graphwalker_amd/build.py) -- results checked against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from gwfiles import write_graph

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "graphwalker_amd", "bin")
COMPAT = os.path.join(ROOT, "graphwalker_amd", "bin", "compat")
KINDS = {"jump": O.JUMP, "stop": O.STOP, "prob": O.PROB, "joint": O.RESTART_JOINT, "restart": O.RESTART}


@pytest.fixture(scope="module")
def graph(tmp_path_factory):
    d = tmp_path_factory.mktemp("g")
    (d / "conf").mkdir()
    (d / "conf" / "graphwalker.cnf").write_text("# empty\n")
    off, col = O.rmat_graph(12, 1 << 16, seed=7)
    base = str(d / "rmat12")
    write_graph(base, off, col, 16)
    write_graph(base, off, col, 1024)
    return d, base, off, col


def run(cmd, cwd, timeout=120, env=None, both=False):
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd, timeout=timeout, env=env)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-3000:])
    return r.stdout + r.stderr if both else r.stdout


@pytest.mark.parametrize("prog", sorted(KINDS))
@pytest.mark.parametrize("schedule", ["persistent", "blocked"])
def test_walkapp_counts_bit_exact(graph, prog, schedule):
    d, base, off, col = graph
    n = off.shape[0] - 1
    R, L, src0 = 5000, 11, 7
    out = str(d / f"counts_{prog}_{schedule}.u32")
    stdout = run([os.path.join(BIN, "walkapp"), "file", base, "program", prog, "R", str(R), "L", str(L),
                  "source", str(src0), "blocksize_kb", "16", "seed", "5", "schedule", schedule, "out", out], d)
    counts = np.fromfile(out, dtype=np.uint32)
    # walk ids follow the packed records' sorted order: by start vertex
    starts = np.sort((src0 + np.arange(R)) % n).astype(np.uint32)
    walks = O.make_walks(starts, sources_=starts, ids=np.arange(R))
    ref = O.replay(O.program(KINDS[prog], L, 0.15, n, seed=5), off, col, walks)
    assert np.array_equal(counts, ref["counts"])
    assert f"walk_visits {ref['visits']}" in stdout


@pytest.mark.parametrize("prog", ["restart", "jump"])
@pytest.mark.parametrize("nmblocks", [1, 3])
def test_walkapp_out_of_core(graph, prog, nmblocks):
    # outofcore 1: the CSR stays in pinned host memory and runs step it
    # through a device cache of nmblocks blocks (the reference's memory budget)
    d, base, off, col = graph
    n = off.shape[0] - 1
    R, L, src0 = 4000, 10, 3
    out = str(d / f"counts_ooc_{prog}_{nmblocks}.u32")
    run([os.path.join(BIN, "walkapp"), "file", base, "program", prog, "R", str(R), "L", str(L), "source", str(src0),
         "blocksize_kb", "16", "nmblocks", str(nmblocks), "outofcore", "1", "seed", "8", "out", out], d)
    counts = np.fromfile(out, dtype=np.uint32)
    starts = np.sort((src0 + np.arange(R)) % n).astype(np.uint32)
    walks = O.make_walks(starts, sources_=starts, ids=np.arange(R))
    ref = O.replay(O.program(KINDS[prog], L, 0.15, n, seed=8), off, col, walks)
    assert np.array_equal(counts, ref["counts"])


@pytest.mark.skipif(not os.access(os.path.join(COMPAT, "simrank"), os.X_OK), reason="compat apps not built")
def test_reference_simrank_unchanged(graph):
    """apps/simrank.cpp of the reference, compiled unchanged against the
    front end, prints the SimRank the oracle's trajectories give."""
    d, base, off, col = graph
    a, b, R, L = 3, 9, 60, 11
    stdout = run([os.path.join(COMPAT, "simrank"), "file", base, "a", str(a), "b", str(b), "R", str(R), "L", str(L),
                  "blocksize_kb", "1024"], d)
    m = re.search(r"SimRank for (\d+) and (\d+) = (\S+)", stdout)
    assert m, stdout[-2000:]
    # records encode(a, cur, i*L) then encode(b, ...): sorted by source, then hop
    prog = O.program(O.RESTART_JOINT, L, 0.15, off.shape[0] - 1, seed=1)
    paths = {}
    for side, s in enumerate((a, b)):
        arr = np.full(R * L, 0xFFFFFFFF, dtype=np.uint64)
        for i in range(R):
            w = O.make_walks([s], sources_=[s], hops=[i * L], ids=[side * R + i])[0]
            tv, th, _ = O.trace(prog, off, col, w)
            arr[th] = tv
        paths[s] = arr
    sim = np.float32(0)
    A, B = paths[a], paths[b]
    for i in range(R):
        for j in range(R):
            for l in range(L):
                if A[i * L + l] == B[j * L + l] and A[i * L + l] != 0xFFFFFFFF:
                    sim = np.float32(np.float64(sim) + (1.0 / (R * R)) * 0.8 ** l)
                    break
    assert m.group(3) == f"{float(sim):g}", (m.group(3), float(sim))


@pytest.mark.skipif(not os.access(os.path.join(COMPAT, "rawrandomwalks"), os.X_OK), reason="compat apps not built")
def test_reference_rawrandomwalks_unchanged(graph):
    d, base, off, col = graph
    stdout = run([os.path.join(COMPAT, "rawrandomwalks"), "file", base, "N", str(off.shape[0] - 1), "R", "20000",
                  "L", "10", "blocksize_kb", "16"], d)
    m = re.search(r"\[walk_steps\] (\d+)", stdout)
    assert m and int(m.group(1)) == 20000 * 10, stdout[-2000:]


@pytest.mark.skipif(not os.access(os.path.join(COMPAT, "graphlet"), os.X_OK), reason="compat apps not built")
@pytest.mark.parametrize("recorder", ["events", "last"])
def test_reference_graphlet_unchanged(tmp_path, recorder):
    """apps/graphlet.cpp of the reference, compiled unchanged against the
    front end: its closed-walk ratio (rand() sources, so distributional) within
    5 sigma of the FP64 probability averaged over uniform sources -- with every
    visit replayed, and with only each walk's last visit recorded on the device
    and replayed (the hop L-1 visit graphlet's updateInfo counts)."""
    from oracle import fp64
    (tmp_path / "conf").mkdir()
    (tmp_path / "conf" / "graphwalker.cnf").write_text("# empty\n")
    off, col = O.rmat_graph(6, 1 << 10, seed=4)
    n = off.shape[0] - 1
    base = str(tmp_path / "rmat6")
    write_graph(base, off, col, 1024)
    R, L = 200_000, 4
    stdout = run([os.path.join(COMPAT, "graphlet"), "file", base, "N", str(n), "R", str(R), "L", str(L),
                  "blocksize_kb", "1024", "recorder", recorder], tmp_path)
    m = re.search(r"cnt_ok:\s+(\d+)", stdout)
    assert m, stdout[-2000:]
    f = float(fp64.return_at(off, col, L, 0.15).mean())
    got = int(m.group(1)) / R
    assert abs(got - f) <= 5 * np.sqrt(f * (1 - f) / R), (got, f)


@pytest.mark.skipif(not os.access(os.path.join(COMPAT, "rwdomination"), os.X_OK), reason="compat apps not built")
def test_reference_rwdomination_unchanged(graph):
    """apps/rwdomination.cpp unchanged: R jump walks of L steps from every
    vertex, N x R x L steps exactly."""
    d, base, off, col = graph
    n = off.shape[0] - 1
    R, L = 3, 6
    stdout = run([os.path.join(COMPAT, "rwdomination"), "file", base, "N", str(n), "R", str(R), "L", str(L),
                  "blocksize_kb", "16"], d)
    m = re.search(r"\[walk_steps\] (\d+)", stdout)
    assert m and int(m.group(1)) == n * R * L, stdout[-2000:]


@pytest.mark.skipif(not os.access(os.path.join(COMPAT, "kk_ppr"), os.X_OK), reason="compat apps not built")
def test_reference_kk_ppr_unchanged(graph):
    """apps/kk_ppr.cpp unchanged: walkspersource uncapped (prob) walks from
    each of numsources consecutive sources; the step total is the oracle's
    for the same walks (source index = field, ids in the packed records'
    sorted order, seed 1)."""
    d, base, off, col = graph
    n = off.shape[0] - 1
    first, ns, wps = 100, 40, 250
    stdout = run([os.path.join(COMPAT, "kk_ppr"), "file", base, "firstsource", str(first), "numsources", str(ns),
                  "walkspersource", str(wps), "maxwalklength", "30", "blocksize_kb", "1024"], d)
    steps = re.search(r"\[walk_steps\] (\d+)", stdout)
    assert steps, stdout[-2000:]
    k = np.repeat(np.arange(ns, dtype=np.uint32), wps)
    walks = O.make_walks(first + k, sources_=k, ids=np.arange(ns * wps))
    ref = O.replay(O.program(O.PROB, 30, 0.15, n, seed=1), off, col, walks)
    assert int(steps.group(1)) == ref["steps"]


@pytest.mark.skipif(not os.access(os.path.join(COMPAT, "msppr"), os.X_OK), reason="compat apps not built")
@pytest.mark.parametrize("recorder,threads", [("events", 1), ("events", 4), ("source", 4)])
def test_reference_msppr_unchanged(graph, tmp_path, recorder, threads):
    """apps/msppr.cpp unchanged (its `killall top` answered by a no-op on
    PATH): the top-20 of source 0's visit distribution it logs equals the
    oracle's counts for the same walks (index 0's walks: ids 0..W-1, seed 1),
    exact because the source reaches fewer vertices than the recorder's
    1000-entry capacity; with every recorder path of the front end."""
    d, base, off, col = graph
    n = off.shape[0] - 1
    wps, L = 600, 10
    prog = O.program(O.STOP, L, 0.15, n, seed=1)
    first, ref = None, None
    for v in range(100, 4000, 37):
        walks = O.make_walks(np.full(wps, v, np.uint32), sources_=np.zeros(wps, np.uint32), ids=np.arange(wps))
        c = O.replay(prog, off, col, walks)["counts"]
        if 40 < np.count_nonzero(c) < 900:
            first, ref = v, c
            break
    assert first is not None
    bindir = tmp_path / "bin"
    bindir.mkdir()
    (bindir / "killall").write_text("#!/bin/sh\nexit 0\n")
    (bindir / "killall").chmod(0o755)
    env = dict(os.environ, PATH=f"{bindir}:{os.environ.get('PATH', '')}")
    out = run([os.path.join(COMPAT, "msppr"), "file", base, "firstsource", str(first), "numsources", "2",
               "walkspersource", str(wps), "maxwalklength", str(L), "blocksize_kb", "1024", "recorder", recorder,
               "replaythreads", str(threads), "execthreads", "4"], d, env=env, both=True)
    rows = re.findall(r"(\d+)-\t(\d+):\t (\d+)", out)
    assert len(rows) == 20, out[-3000:]
    assert [int(r[0]) for r in rows] == list(range(20))
    counts = [int(r[2]) for r in rows]
    assert counts == sorted(counts)  # the reference's order: fewest visits first
    for _, vtx, cnt in rows:
        assert ref[int(vtx)] == int(cnt), (vtx, cnt, ref[int(vtx)])
    assert sorted(counts) == sorted(np.sort(ref)[-20:].tolist())


@pytest.mark.skipif(not os.access(os.path.join(COMPAT, "rwdomination"), os.X_OK), reason="compat apps not built")
def test_reference_rwdomination_vertex_recorder(graph):
    """rwdomination unchanged with the device per-vertex recorder: the same
    step total as the event-log path."""
    d, base, off, col = graph
    n = off.shape[0] - 1
    outs = []
    for rec in ("events", "vertex"):
        stdout = run([os.path.join(COMPAT, "rwdomination"), "file", base, "N", str(n), "R", "2", "L", "6",
                      "blocksize_kb", "16", "recorder", rec, "replaythreads", "4"], d)
        m = re.search(r"\[walk_steps\] (\d+)", stdout)
        assert m, stdout[-2000:]
        outs.append(int(m.group(1)))
    assert outs[0] == outs[1] == n * 2 * 6


@pytest.mark.skipif(not os.access(os.path.join(COMPAT, "rawrandomwalks"), os.X_OK), reason="compat apps not built")
def test_reference_rawrandomwalks_no_recorder(graph):
    d, base, off, col = graph
    stdout = run([os.path.join(COMPAT, "rawrandomwalks"), "file", base, "N", str(off.shape[0] - 1), "R", "20000",
                  "L", "10", "blocksize_kb", "16", "recorder", "none"], d)
    m = re.search(r"\[walk_steps\] (\d+)", stdout)
    assert m and int(m.group(1)) == 20000 * 10, stdout[-2000:]
