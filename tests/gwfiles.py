"""Writes graphs in the GraphWalker preprocessed file format for app tests
(<g>_GraphWalker/graphinfo/file_0.{beg_pos,csr}, N.nvertices,
filesize_1024GB.filerange, blocksize_<KB>KB.blockrange)."""
import os

import numpy as np


def block_table(off: np.ndarray, blocksize_kb: int) -> list[int]:
    """The reference's block cut rule (preprocess/conversions.hpp:300-364)."""
    mneb = blocksize_kb * 1024 // 4
    blocks = [0]
    start = 0
    off = [int(x) for x in off]
    for v in range(1, len(off)):
        if off[v] - start > mneb:
            if off[v] - off[v - 1] > mneb:
                blocks += [v - 1, v]
                start = off[v]
            else:
                blocks.append(v - 1)
                start = off[v - 1]
    blocks.append(len(off) - 1)
    return blocks


def write_graph(base: str, off: np.ndarray, col: np.ndarray, blocksize_kb: int) -> int:
    d = base + "_GraphWalker"
    os.makedirs(os.path.join(d, "graphinfo"), exist_ok=True)
    np.asarray(off, dtype=np.uint64).tofile(os.path.join(d, "graphinfo", "file_0.beg_pos"))
    np.asarray(col, dtype=np.uint32).tofile(os.path.join(d, "graphinfo", "file_0.csr"))
    n = len(off) - 1
    with open(os.path.join(d, "filesize_1024GB.filerange"), "w") as f:
        f.write(f"0\n{n}\n")
    with open(os.path.join(d, "N.nvertices"), "w") as f:
        f.write(f"{n}\n")
    blocks = block_table(off, blocksize_kb)
    with open(os.path.join(d, f"blocksize_{blocksize_kb}KB.blockrange"), "w") as f:
        f.write("".join(f"{b}\n" for b in blocks))
    return len(blocks) - 1
