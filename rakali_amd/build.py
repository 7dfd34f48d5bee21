"""Build recipe for the in-tree libraries ``rakali_amd/librakali_amd.so``
(the HIP step, ``rkg_*``) and ``rakali_amd/librakali.so`` (the handle API,
``rki_*``, host code linked against the former).

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
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "librakali_amd.so"
SOURCES = [CSRC / "rkg_api.hip", CSRC / "rkg_ensemble.hip", CSRC / "rkg_adjoint.hip", CSRC / "rkg_gauges.hip", CSRC / "rkg_noalloc.hip",
           *(CSRC / f"stage_feat{k}.hip" for k in range(8))]
HEADERS = [CSRC / "swe_core.hpp", CSRC / "swe_kernels.hpp", CSRC / "crmath.hpp", CSRC / "adjoint.hpp",
           CSRC / "stage_kernels.hpp", CSRC / "stage_launch.hpp", CSRC / "stage_feat.inc",
           CSRC / "rkg_internal.hpp", ROOT / "include" / "rakali_amd.h"]
RKI_LIB = PKG / "librakali.so"
RKI_SOURCES = [CSRC / "rki_api.cpp"]
RKI_HEADERS = [CSRC / "crmath.hpp", CSRC / "nc_classic.hpp", ROOT / "include" / "rakali.h",
               ROOT / "include" / "rakali_amd.h"]
DRIVER = PKG / "rakali_driver"
DRIVER_SOURCES = [PKG / "fortran" / "rakali_amd_api.F90", PKG / "fortran" / "rakali_driver.F90"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; ROCm is required to build rakali_amd")


def flang() -> str | None:
    for cand in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"),
                 shutil.which("flang")):
        if cand and Path(cand).exists():
            return cand
    return None


def build_driver(verbose: bool = False) -> Path | None:
    """The Fortran namelist driver (iso_c_binding over librakali.so), or
    None when no Fortran compiler is installed."""
    fc = flang()
    if fc is None:
        return None
    import tempfile
    with tempfile.TemporaryDirectory() as mod_dir:
        tmp = DRIVER.with_suffix(".tmp")
        cmd = [fc, "-O2", "-module-dir", mod_dir, "-o", str(tmp), *map(str, DRIVER_SOURCES),
               "-L", str(PKG), "-l:librakali.so", "-l:librakali_amd.so", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"{fc} failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
        os.replace(tmp, DRIVER)
    return DRIVER


def compile_flags() -> list[str]:
    return [
        "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(ROOT / "include"),
    ]


def _run(cmd: list[str], out: Path, verbose: bool) -> Path:
    tmp = out.with_suffix(".so.tmp")
    cmd = [*cmd, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


def unit_flags(src: Path) -> list[str]:
    """Extra flags of one translation unit: the stage kernels (stage_feat*)
    use LLVM's ILP-first machine scheduler, which is about 1 % faster on the
    4096x4096 step at the same occupancy (three waves per SIMD, no spills);
    it changes the instruction order, not the arithmetic."""
    if src.stem.startswith("stage_feat"):
        return ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
    return []


def _compile_objects(sources: list[Path], outdir: Path, verbose: bool) -> list[str]:
    """Compiles each translation unit to an object, the units in parallel
    (each one instantiates many stage kernels and takes tens of seconds)."""
    flags = [f for f in compile_flags() if f != "-shared"]

    def one(src: Path) -> str:
        obj = outdir / (src.stem + ".o")
        cmd = [hipcc(), *flags, *unit_flags(src), "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name} ({res.returncode}):\n{res.stdout}\n{res.stderr}")
        return str(obj)

    jobs = max(1, min(len(sources), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as pool:
        return list(pool.map(one, sources))


def build(verbose: bool = False) -> Path:
    """Compile both libraries unconditionally; return the step library's path."""
    with tempfile.TemporaryDirectory(prefix="rakali_amd_build_") as tmp:
        objs = _compile_objects(SOURCES, Path(tmp), verbose)
        _run([hipcc(), *compile_flags(), *objs], LIB, verbose)
    # The handle API is host code (it uses the step's host-side sin/cos);
    # it links the step library by soname and finds it next to itself.
    _run([hipcc(), *compile_flags(), "-x", "hip", *map(str, RKI_SOURCES),
          "-L", str(PKG), "-l:librakali_amd.so", "-Wl,-rpath,$ORIGIN",
          "-I", str(CSRC)], RKI_LIB, verbose)
    build_driver(verbose)
    return LIB


def is_stale() -> bool:
    for lib, srcs in ((LIB, SOURCES + HEADERS), (RKI_LIB, RKI_SOURCES + RKI_HEADERS + [LIB])):
        if not lib.exists():
            return True
        t = lib.stat().st_mtime
        if any(p.exists() and p.stat().st_mtime > t for p in srcs):
            return True
    return False


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
