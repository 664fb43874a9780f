"""Writes tests/golden/philox_vectors.json: Philox4x32-10 outputs computed by
rocRAND's header implementation (compiled here for the host with hipcc) on
the three Random123 known-answer inputs plus 61 seeded random counter/key
pairs.  Run in a container with ROCm headers:  python tests/golden/make_philox_vectors.py
"""
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> None:
    rng = np.random.default_rng(20190712)
    inputs = [
        [0, 0, 0, 0, 0, 0],
        [0xFFFFFFFF] * 6,
        [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344, 0xA4093822, 0x299F31D0],
    ]
    inputs += rng.integers(0, 2**32, size=(61, 6), dtype=np.uint64).tolist()
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "check")
        subprocess.run(["hipcc", "-O2", "-o", exe, os.path.join(HERE, "philox_rocrand_check.cpp")], check=True)
        text = "\n".join(" ".join(f"{int(x):08x}" for x in row) for row in inputs) + "\n"
        out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    vectors = []
    for row, line in zip(inputs, out):
        o = [int(x, 16) for x in line.split()]
        vectors.append({"ctr": [int(x) for x in row[:4]], "key": [int(x) for x in row[4:]], "out": o})
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump({"source": "rocRAND rocrand_philox4x32_10.h ten_rounds (host build)", "vectors": vectors}, f, indent=0)


if __name__ == "__main__":
    main()
