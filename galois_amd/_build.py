"""Build recipe for libgalois_amd.so (HIP kernels for gfx950 + C++ host code).

Every source under galois_amd/csrc is compiled with hipcc into an object under
galois_amd/_build/ and linked into galois_amd/libgalois_amd.so.  The build is
in-tree so the shared library travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libgalois_amd.so"
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-but-set-variable"]
# Packed FP32 instructions are disabled for all device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def sources() -> list[Path]:
    return sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp")])


def build_library(jobs: int | None = None, verbose: bool = False) -> Path:
    """Compile every source and link libgalois_amd.so; returns the library path."""
    OBJ.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    build_apps(jobs)
    return LIB


APPS = PKG / "apps"
BIN = PKG / "bin"
# Lonestar app entry points: source stem -> executable name of the reference.
APP_NAMES = {"bfs": "bfs", "sssp": "sssp", "connected_components": "connected-components", "kcore": "k-core",
             "pagerank_pull": "pagerank-pull", "betweennesscentrality": "betweennesscentrality",
             "graph_convert": "graph-convert", "pagerank_push": "pagerank-push",
             "bfs_directionopt": "bfs-directionopt"}


def _build_app(src: Path) -> Path:
    exe = BIN / APP_NAMES[src.stem]
    cmd = [HIPCC, *COMMON, "-I", str(PKG.parent / "include"), str(src), "-o", str(exe), f"-L{PKG}", "-lgalois_amd",
           "-Wl,-rpath,$ORIGIN/.."]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"app build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return exe


def build_apps(jobs: int | None = None) -> list[Path]:
    """Build the lonestar-compatible app executables into galois_amd/bin/."""
    BIN.mkdir(exist_ok=True)
    srcs = sorted(APPS.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=jobs or len(srcs) or 1) as ex:
        return list(ex.map(_build_app, srcs))


if __name__ == "__main__":
    build_library(verbose=True)
