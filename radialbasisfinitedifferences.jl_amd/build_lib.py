"""Build librbffd.so (hand-written gfx950 HIP kernels + the C ABI) in-tree.

Every translation unit under csrc/ is compiled by hipcc for gfx950 only, in
parallel, then linked into ``rbffd_amd/librbffd.so`` next to the Python
mirror that loads it.  Packed FP32 instructions are disabled for all device
code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
OUT = os.path.join(PKG, "rbffd_amd", "librbffd.so")
SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the local-solve kernels unroll their elimination loops completely
    "-mllvm", "-pragma-unroll-threshold=2000000",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _stale(src: str, obj: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps.append(os.path.join(ROOT, "include", "rbffd.h"))
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, force: bool = True) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    if not force and not _stale(src, obj):
        return obj
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, force: bool = True) -> str:
    """Compile every source (force) or only the out-of-date ones, then link."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(SOURCES)))) as ex:
        objs = list(ex.map(lambda f: _compile(f, force), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True, force="--incremental" not in sys.argv)
    sys.exit(0)
