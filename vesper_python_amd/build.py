"""Build libvesper_hip.so (hand-written HIP kernels for gfx950 + the C-ABI).

Usage:  python -m vesper_python_amd.build [--incremental]   (or __graft_entry__.build())

The library is built in-tree (vesper_python_amd/libvesper_hip.so) with
explicit hipcc invocations: one object per translation unit, then a link.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libvesper_hip.so")
SOURCES = ["vx_search.hip", "vx_slab.hip", "vx_fastvec.hip", "vx_score.hip", "vx_mlp.hip", "vx_interp.hip",
           "vx_noalloc.hip", "vx_rescore.hip", "vx_propose.hip", "vx_xmr.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VX_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the parity-critical float arithmetic (rotation positions,
# fastVEC geometry, the prob-mode mix) is written with explicit fma / rounded
# intrinsics and must not be re-associated or contracted.
# -packed-fp32-ops off: keep the compiler from emitting packed FP32 math.
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


# Per-source additions: the radix-split MFMA x pass keeps its accumulators in
# ordinary VGPRs (no AGPR copies), which leaves it 3 waves per SIMD at N = 175.
EXTRA_FLAGS = {"vx_xmr.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _stale(src: str, obj: str) -> bool:
    """True unless obj is newer than src and every header it can include."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps.append(os.path.join(INCLUDE, "vesper_hip.h"))
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) >= t for d in deps)


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [HIPCC, *CFLAGS, *EXTRA_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True, incremental: bool = False) -> str:
    """Compile every source (incremental=True: only those older than their
    object, for development) and link the library."""
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda f: _compile(f, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
