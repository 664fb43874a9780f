"""Every file:line citation of the reference in this repository's sources and
documents names lines that exist in that reference file (where the
reference tree is present; the GPU box has none and skips this)."""
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree not present")

CITE = re.compile(r"\b([A-Za-z_][A-Za-z0-9_]*\.(?:hpp|cpp|cu|h|statistics))[:](\d+)(?:-(\d+))?")


def _sources():
    pats = ["include/*.h", "graphwalker_amd/csrc/*", "graphwalker_amd/cxx/**/*", "graphwalker_amd/apps/*",
            "graphwalker_amd/*.py", "oracle/*", "tests/*.py", "bench.py", "__graft_entry__.py", "DESIGN.md",
            "INTEGRATION.md", "README.md"]
    for p in pats:
        for f in glob.glob(os.path.join(ROOT, p), recursive=True):
            if os.path.isfile(f) and not f.endswith((".so", ".o", ".pyc")) and os.path.basename(f) != "test_citations.py":
                yield f


def _ref_lengths():
    lengths = {}
    for dirpath, _, files in os.walk(REF):
        for f in files:
            path = os.path.join(dirpath, f)
            try:
                with open(path, "rb") as fh:
                    n = fh.read().count(b"\n") + 1
            except OSError:
                continue
            lengths.setdefault(f, []).append(n)
    return lengths


def test_reference_citations_name_existing_lines():
    lengths = _ref_lengths()
    bad, seen = [], 0
    for f in _sources():
        try:
            text = open(f, encoding="utf-8").read()
        except (UnicodeDecodeError, OSError):
            continue
        for m in CITE.finditer(text):
            name, a, b = m.group(1), int(m.group(2)), int(m.group(3) or m.group(2))
            if name not in lengths:
                continue  # a file of this repository, not the reference's
            seen += 1
            if not (1 <= a <= b and any(b <= n for n in lengths[name])):
                line = text[: m.start()].count("\n") + 1
                bad.append(f"{os.path.relpath(f, ROOT)}:{line}: {m.group(0)} (file has {max(lengths[name])} lines)")
    assert seen > 50
    assert not bad, "\n".join(bad)
