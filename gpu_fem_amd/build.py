"""Build libgpufem_amd.so (gfx950) in-tree with hipcc.

Objects go to ``_build/`` at the repository root; the shared library lands
next to this file so it travels with the source tree.  Every call compiles
(incrementally by timestamp unless ``force``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build", "gpufem_amd")
LIB = os.path.join(HERE, "libgpufem_amd.so")
SOURCES = ["symbolic.cpp", "amd.cpp", "bdf.cpp", "vtk.cpp", "capi.cpp", "server.cpp", "assembly.hip", "bsr.hip",
           "ilu.hip", "krylov.hip"]
HEADERS = ["common.h", "symbolic.h", "krylov.h", "mitc4.h", "problem.h"]
ARCH = os.environ.get("GPUFEM_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-munsafe-fp-atomics",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _newest_header():
    hs = [os.path.join(CSRC, h) for h in HEADERS]
    hs.append(os.path.join(ROOT, "include", "gpufem_amd.h"))
    return max(os.path.getmtime(h) for h in hs)


def _compile(src, force):
    obj = os.path.join(BUILD, src + ".o")
    path = os.path.join(CSRC, src)
    if (not force and os.path.exists(obj) and
            os.path.getmtime(obj) >= max(os.path.getmtime(path),
                                         _newest_header())):
        return obj
    cmd = [HIPCC] + FLAGS + ["-c", path, "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(force: bool = False, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread",
           "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_apps()
    if verbose:
        print(f"built {LIB}")
    return LIB


APPS = ["linear_static"]
BIN = os.path.join(HERE, "bin")


def _build_apps():
    """C++ programs written against include/gpufem/gpufem.hpp."""
    os.makedirs(BIN, exist_ok=True)
    for app in APPS:
        src = os.path.join(HERE, "apps", app + ".cpp")
        out = os.path.join(BIN, app)
        cmd = [HIPCC, "-O2", "-std=c++17", "-o", out, src, "-L" + HERE,
               "-lgpufem_amd", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"build of {app} failed:\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
