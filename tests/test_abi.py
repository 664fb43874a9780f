"""The C-ABI library loads and exports every entry point include/gwhip.h
declares; host-side argument checks that run without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import graphwalker_amd as gw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gwhip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gw_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "gw_run" in syms and "gw_create" in syms and len(syms) >= 15


def test_library_exports_every_declared_symbol():
    assert os.path.exists(gw.LIB_PATH), "build the library first (python -m graphwalker_amd.build)"
    lib = ctypes.CDLL(gw.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    nm = subprocess.run(["nm", "-D", "--defined-only", gw.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (gw_[a-z_]+)\b", nm))
    assert set(declared_symbols()) <= exported


def test_python_binding_covers_header():
    L = gw.lib()
    for s in declared_symbols():
        assert getattr(L, s).argtypes is not None or s == "gw_destroy", s


def test_header_compiles_as_c():
    src = '#include "gwhip.h"\nint main(void){ gw_ctx *c = 0; return gw_create(&c, 0) == 12345; }\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        "-x", "c", "-"], input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_header():
    assert ctypes.sizeof(gw.Program) == 72
    assert gw.WALK_DTYPE.itemsize == 16
    src = ('#include "gwhip.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(void){printf("%zu %zu %zu %zu",'
           ' sizeof(gw_program), sizeof(gw_walk), sizeof(gw_stats), offsetof(gw_program, recorder)); return 0;}\n')
    exe = "/tmp/gw_layout_probe"
    r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-x", "c", "-", "-o", exe],
                       input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == ctypes.sizeof(gw.Program)
    assert int(out[1]) == gw.WALK_DTYPE.itemsize
    assert int(out[2]) == ctypes.sizeof(gw.Stats)
    assert int(out[3]) == gw.Program.recorder.offset


def test_stop_threshold():
    assert gw.stop_threshold(0.0) == 0
    assert gw.stop_threshold(1.0) == 1 << 32
    assert gw.stop_threshold(0.15) == 644245094
    with pytest.raises(ValueError):
        gw.stop_threshold(1.5)


def test_program_names():
    p = gw.program("stop", 20, 0.15, seed=3, schedule="persistent")
    assert p.kind == gw.STOP and p.schedule == gw.SCHED_PERSISTENT and p.L == 20


def test_engine_without_gpu_fails_loudly():
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(gw.GwError):
        gw.Engine()


def test_null_context_is_rejected():
    L = gw.lib()
    assert L.gw_run(None, None, None) != 0
    assert L.gw_graph_set(None, None, None, 0, 0) != 0
    assert L.gw_last_error(None) == b"null context"


def test_sources_match_oracle():
    import oracle as O

    for seed, n, k in ((3, 1 << 26, 4096), (1, 1000, 77), (2**40 + 5, 7, 10)):
        assert np.array_equal(gw.sources(seed, n, k), O.sources(seed, n, k))


def test_python_constants_match_header():
    """Every GW_<NAME> integer #define the binding mirrors as <NAME> has the header's value."""
    defs = dict((k, int(v, 0)) for k, v in re.findall(r"#define GW_([A-Z0-9_]+)\s+(0x[0-9a-fA-F]+|\d+)u?\b",
                                                       open(HEADER).read()))
    mirrored = [k for k in defs if hasattr(gw, k) and isinstance(getattr(gw, k), int)]
    assert {"REC_EVENTS", "REC_EVENTS_LAST", "REC_NO_FINAL"} <= set(mirrored)
    for k in mirrored:
        assert getattr(gw, k) == defs[k], k
