"""Build recipe for the HIP extension (gfx950 only).

Compiles every ``*.hip`` file in this directory with hipcc into objects under
``mr_recon_amd/_build/`` and links them into ``mr_recon_amd/libmr_recon_amd.so``.
The library exports only the C-ABI declared in ``include/mr_recon_amd.h`` and
links against the HIP runtime and hipFFT (rocFFT); PyTorch loads it with ctypes.
In a PyTorch process the hipFFT already loaded by torch (same soname) is reused.

Usage: ``python -m mr_recon_amd.csrc.build`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG = CSRC.parent
BUILD = PKG / "_build"
LIB = PKG / "libmr_recon_amd.so"

ARCH = "gfx950"
CXXFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # Packed FP32 VALU ops are disabled for every kernel in this library.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]
# Per-file extra flags. Binning reproduces the reference's separately rounded
# fp32 multiply and add bit for bit, so no FMA contraction there.
# bin.hip: coordinate arithmetic must round like separate fp32 operations.
# nufft_kernels.hip: matrix-core accumulators in ordinary VGPRs (the gridding kernel reads
# them with vector ALU ops at every cell end; in AGPRs each group would copy them out).
FILE_FLAGS = {"bin.hip": ["-ffp-contract=off"], "nufft_kernels.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def _compile(src: Path, verbose: bool) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [hipcc(), *CXXFLAGS, *FILE_FLAGS.get(src.name, []), "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile all sources and link the shared library; returns its path."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_name(LIB.name + ".tmp")
    rocm_lib = Path(hipcc()).resolve().parent.parent / "lib"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), f"-L{rocm_lib}", "-lhipfft",
           "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
