"""Build libpsl_amd.so (HIP, gfx950) in-tree.

Each .hip translation unit is compiled with hipcc for gfx950 and linked into
periodicshiftlbm_amd/libpsl_amd.so, which exports the C ABI of include/psl.h.
Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpsl_amd.so")

ARCH = os.environ.get("PSL_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-I", os.path.join(REPO, "include")]
# per-file extra flags
EXTRA = {
    # voxelizer masks must be bit-identical to the CPU restatement
    "psl_stl.hip": ["-ffp-contract=off"],
    # DIAD builder: marker-rule positions and shifted origins round as written
    "psl_diad_build.hip": ["-ffp-contract=off"],
}
SOURCES = ["psl_grid.hip", "psl_refine.hip", "psl_stl.hip", "psl_esotwist.hip", "psl_batch.hip", "psl_closure.hip",
           "psl_diad_build.hip", "psl_output.hip",
           "psl_scalar.hip", "psl_probe.hip", "psl_state.hip"]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libpsl_amd.so")
    return p


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [hipcc(), *COMMON, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


EXAMPLES = os.path.join(HERE, "examples")
EXAMPLE_BIN = os.path.join(EXAMPLES, "bin")


def example_sources():
    return sorted(f for f in os.listdir(EXAMPLES) if f.endswith(".cpp"))


def build_example(src: str, out_dir: str = EXAMPLE_BIN) -> str:
    """One host driver of examples/ against include/psl_facade.hpp and
    libpsl_amd.so (g++ only: the facade is plain C++ over the C ABI).  The
    rpath is relative to the binary, so the tree can move."""
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, src[:-4])
    rpath = os.path.relpath(HERE, out_dir)
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-I", os.path.join(REPO, "include"),
           os.path.join(EXAMPLES, src), "-L", HERE, "-lpsl_amd", f"-Wl,-rpath,$ORIGIN/{rpath}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"example build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return exe


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for src in example_sources():
        build_example(src)
    if verbose:
        print(f"built {LIB} and {len(example_sources())} example drivers in {EXAMPLE_BIN}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
