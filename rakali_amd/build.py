"""Build recipe for the in-tree HIP library ``rakali_amd/librakali_amd.so``.

The library is compiled for gfx950 only.  Device code is built with
``-ffp-contract=off`` so that no multiply and add are fused: every FP64
operation of the step rounds exactly as the reference's x86-64 build does.
Packed FP32 instructions are disabled (the step is FP64 throughout; this
keeps the compiler from emitting them anywhere in the library).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "librakali_amd.so"
SOURCES = [CSRC / "rkg_api.hip"]
HEADERS = [CSRC / "swe_core.hpp", CSRC / "swe_kernels.hpp", CSRC / "crmath.hpp",
           ROOT / "include" / "rakali_amd.h"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; ROCm is required to build rakali_amd")


def compile_flags() -> list[str]:
    return [
        "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(ROOT / "include"),
    ]


def build(verbose: bool = False) -> Path:
    """Compile the library unconditionally and return its path."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *compile_flags(), "-o", str(tmp), *map(str, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def is_stale() -> bool:
    if not LIB.exists():
        return True
    t = LIB.stat().st_mtime
    return any(p.stat().st_mtime > t for p in SOURCES + HEADERS)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
