"""Builds libgwhip.so (HIP, gfx950) in-tree: graphwalker_amd/lib/libgwhip.so.

    python -m graphwalker_amd.build

Each ``csrc/*.hip`` unit is compiled separately (in parallel) and linked with
hipcc.  Packed-FP32 instruction selection is disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libgwhip.so")
APPS = os.path.join(HERE, "apps")
BIN = os.path.join(HERE, "bin")
CXX_INC = os.path.join(HERE, "cxx")
C_INC = os.path.join(os.path.dirname(HERE), "include")
ARCH = os.environ.get("GW_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o"))
    cmd = [_hipcc(), *HIPCC_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_apps() -> list[str]:
    """Host apps on the GraphWalker API (apps/*.cpp) -> bin/, linked to lib/libgwhip.so."""
    os.makedirs(BIN, exist_ok=True)
    out = []
    for f in sorted(os.listdir(APPS)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(BIN, f[:-4])
        cmd = ["g++", "-std=c++17", "-O2", "-fopenmp", "-Wall", "-I", CXX_INC, "-I", C_INC, os.path.join(APPS, f),
               "-L", LIB_DIR, "-lgwhip", "-Wl,-rpath,$ORIGIN/../lib", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"app build failed: {' '.join(cmd)}\n{r.stderr}")
        out.append(exe)
    return out


REFERENCE_APPS = ("rawrandomwalks", "msppr", "simrank", "graphlet", "rwdomination", "kk_ppr")


def build_compat_apps(reference: str = "/root/reference") -> list[str]:
    """The reference's own apps, compiled unchanged from where they lie
    against this package's GraphWalker API headers (cxx/) and linked to
    lib/libgwhip.so -> bin/compat/ (the drop-in boundary).  Skipped where the
    reference tree is absent; nothing of it is copied."""
    apps_dir = os.path.join(reference, "apps")
    if not os.path.isdir(apps_dir):
        return []
    out_dir = os.path.join(BIN, "compat")
    os.makedirs(out_dir, exist_ok=True)
    out = []
    for app in REFERENCE_APPS:
        src = os.path.join(apps_dir, app + ".cpp")
        if not os.path.exists(src):
            continue
        exe = os.path.join(out_dir, app)
        cmd = ["g++", "-std=c++17", "-O2", "-fopenmp", "-w", "-I", CXX_INC, "-I", C_INC, src, "-L", LIB_DIR, "-lgwhip",
               "-Wl,-rpath,$ORIGIN/../../lib", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compat app build failed: {' '.join(cmd)}\n{r.stderr}")
        out.append(exe)
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    apps = build_apps()
    compat = build_compat_apps()
    if verbose:
        print(f"built {LIB}, {len(apps)} apps and {len(compat)} unchanged reference apps")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
