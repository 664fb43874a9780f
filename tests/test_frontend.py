"""C++ front end (graphwalker_amd/cxx): the reference apps compile unchanged
against it, and its preprocessing writes the reference's file format."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from gwfiles import block_table

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
BIN = os.path.join(ROOT, "graphwalker_amd", "bin")
APPS = ["rawrandomwalks", "msppr", "simrank", "graphlet", "rwdomination", "kk_ppr"]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "apps")), reason="GraphWalker sources not present")
@pytest.mark.parametrize("app", APPS)
def test_reference_app_compiles_unchanged(app):
    cmd = ["g++", "-std=c++17", "-fopenmp", "-fsyntax-only", "-Wno-unused-result",
           "-I", os.path.join(ROOT, "graphwalker_amd", "cxx"), "-I", os.path.join(ROOT, "include"),
           os.path.join(REF, "apps", app + ".cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_front_end_apps_are_built():
    for exe in ("walkapp", "gw_convert"):
        assert os.access(os.path.join(BIN, exe), os.X_OK), f"{exe} not built (python -m graphwalker_amd.build)"


def _edge_list(path, n=300, m=3000, seed=4):
    rng = np.random.default_rng(seed)
    src = np.sort(rng.integers(0, n - 20, m))  # sorted by source, as the reference converter needs
    dst = rng.integers(0, n, m)
    dst[::97] = src[::97]  # some self-loops
    with open(path, "w") as f:
        f.write("# comment\n")
        for s, d in zip(src, dst):
            f.write(f"{s}\t{d}\n")
    return src, dst


def test_gw_convert_format(tmp_path):
    g = str(tmp_path / "g.txt")
    src, dst = _edge_list(g)
    r = subprocess.run([os.path.join(BIN, "gw_convert"), "file", g, "blocksize_kb", "1"], capture_output=True,
                       text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    keep = src != dst
    n = int(max(src.max(), dst.max())) + 1
    off = np.fromfile(g + "_GraphWalker/graphinfo/file_0.beg_pos", dtype=np.uint64)
    col = np.fromfile(g + "_GraphWalker/graphinfo/file_0.csr", dtype=np.uint32)
    exp_off = np.concatenate([[0], np.cumsum(np.bincount(src[keep], minlength=n))])
    assert np.array_equal(off, exp_off)
    assert np.array_equal(col, dst[keep])  # rows keep file order
    blocks = [int(x) for x in open(g + "_GraphWalker/blocksize_1KB.blockrange").read().split()]
    assert blocks == block_table(off, 1)
    assert open(g + "_GraphWalker/N.nvertices").read().split() == [str(n)]


@pytest.mark.skipif(not os.access(os.path.join(ROOT, "oracle", "_ref", "rawrandomwalks"), os.X_OK),
                    reason="reference apps not built (oracle/build_ref.sh)")
def test_conversion_bytes_equal_reference_converter(tmp_path):
    """The reference's own converter (a reference app built from its sources)
    and gw_convert write identical beg_pos / csr / blockrange files."""
    a, b = tmp_path / "ref", tmp_path / "ours"
    a.mkdir()
    b.mkdir()
    (a / "conf").mkdir()
    (a / "conf" / "graphwalker.cnf").write_text("# empty\n")
    _edge_list(str(a / "g.txt"), n=500, m=6000, seed=9)
    shutil.copy(a / "g.txt", b / "g.txt")
    ref = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "rawrandomwalks"), "file", str(a / "g.txt"),
                          "N", "500", "R", "10", "L", "2", "blocksize_kb", "2", "execthreads", "1"],
                         capture_output=True, text=True, cwd=a, timeout=60)
    assert ref.returncode == 0, ref.stderr[-2000:]
    ours = subprocess.run([os.path.join(BIN, "gw_convert"), "file", str(b / "g.txt"), "blocksize_kb", "2"],
                          capture_output=True, text=True, cwd=b)
    assert ours.returncode == 0, ours.stderr
    for rel in ("graphinfo/file_0.beg_pos", "graphinfo/file_0.csr", "blocksize_2KB.blockrange", "N.nvertices"):
        x = (a / "g.txt_GraphWalker" / rel).read_bytes()
        y = (b / "g.txt_GraphWalker" / rel).read_bytes()
        assert x == y, rel


def test_block_table_rule():
    off = np.array([0, 10, 20, 300, 310, 320, 330, 600, 600, 610], dtype=np.uint64)
    # 256 columns per KB: vertex 2 (280 edges) and vertex 6 (270) exceed it
    assert block_table(off, 1) == [0, 2, 3, 6, 7, 9]
    # a single block when everything fits
    assert block_table(off, 4) == [0, 9]
    # the cut rule on an R-MAT graph keeps blocks within the edge budget,
    # except single oversized vertices
    o, _ = O.rmat_graph(12, 1 << 16, seed=3)
    bt = block_table(o, 16)
    for x, y in zip(bt, bt[1:]):
        assert y >= x
        if y - x > 1:
            assert int(o[y]) - int(o[x]) <= 16 * 256 + int(o[y] - o[y - 1])


def test_block_cut_matches_the_reference_rule():
    # the product's vectorised cut against the rule written out vertex by vertex
    import graphwalker_amd as gw
    from gwfiles import block_table

    rng = np.random.default_rng(3)
    for n, kb in ((1, 1), (50, 1), (3000, 1), (3000, 4), (20000, 16)):
        deg = rng.geometric(0.05, n) - 1
        deg[rng.integers(0, n, max(1, n // 500))] += 3000  # rows larger than a block
        off = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
        assert gw.block_cut(off, kb).tolist() == block_table(off, kb), (n, kb)
