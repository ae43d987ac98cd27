"""Compile the HIP sources of mspai_amd into libmspai_gpu.so (gfx950).

hipcc cross-compiles without a GPU.  Flags:
  -ffp-contract=off   no implicit fused multiply-add: the FP64 kernels use
                      explicit __fma_rn wherever the fixed operation order
                      has a fused multiply-add, and nothing else fuses.
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keeps the compiler from emitting packed FP32 ops.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmspai_gpu.so")
CLI = os.path.join(HERE, "mspai_gpu")
CLI_SOURCES = ["host/mspai_cli.cpp", "host/mmio.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["capi.hip", "spai_columns.hip", "apply.hip", "spai_block.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]


def _compile(src: str, obj_dir: str) -> str:
    obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    build_cli()
    if verbose:
        print("built", OUT, "and", CLI)
    return OUT


def build_cli() -> str:
    """The mspai_gpu command-line builder (host C++ over the C ABI)."""
    tmp = CLI + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", *[os.path.join(HERE, s) for s in CLI_SOURCES],
           "-L", HERE, "-lmspai_gpu", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
