"""Build the gfx950 HIP library ``lib/librsdft_gpu.so`` and the Fortran shim.

Objects compile in parallel with ``hipcc --offload-arch=gfx950``; packed
FP32 instructions are disabled for all device code.  The Fortran binding
module (``fortran/rsdft_gpu_mod.f90``) is compiled with ``amdflang`` when it
is on the PATH.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(LIBDIR, "librsdft_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-result",
]
SOURCES = ["capi.cpp", "tables.cpp", "setup.hip", "hpsi_f64.hip", "hpsi_f32.hip", "solver.hip", "scf.hip", "lpc.hip",
           "scf_driver.cpp"]


def _stale(obj: str, src: str) -> bool:
    """True when obj is older than its source or any header it may include."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    deps += [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE)]
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(obj, src):
        return obj
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC, *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every source (or, with incremental, only those whose object is
    older than the source or a header) and link librsdft_gpu.so."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda src: _compile(src, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    """Fortran drop-in module (librsdft_gpu_f.so) and the Fortran SCF driver
    (lib/rsdft_scf), built with AMD flang when it is installed."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    src = os.path.join(HERE, "fortran", "rsdft_gpu_mod.f90")
    prog = os.path.join(HERE, "fortran", "rsdft_scf.f90")
    if fc is None or not os.path.exists(src):
        return None
    out = os.path.join(LIBDIR, "librsdft_gpu_f.so")
    exe = os.path.join(LIBDIR, "rsdft_scf")
    moddir = os.path.join(BUILD, "mod")
    os.makedirs(moddir, exist_ok=True)
    steps = [
        [fc, "-O2", "-fPIC", "-shared", "-Wl,-soname,librsdft_gpu_f.so", "-module-dir", moddir, "-o", out + ".tmp",
         src, f"-L{LIBDIR}", "-lrsdft_gpu", "-Wl,-rpath,$ORIGIN"],
        [fc, "-O2", "-I", moddir, "-o", exe + ".tmp", prog, f"-L{LIBDIR}", "-lrsdft_gpu_f", "-lrsdft_gpu",
         "-Wl,-rpath,$ORIGIN"],
    ]
    r = subprocess.run(steps[0], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran build failed: {' '.join(steps[0])}\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)  # the program links the module library by its final name
    r = subprocess.run(steps[1], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran build failed: {' '.join(steps[1])}\n{r.stdout}\n{r.stderr}")
    os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {out} and {exe}")
    return out


def main() -> None:
    build_library(verbose=True)
    build_fortran(verbose=True)


if __name__ == "__main__":
    sys.exit(main())
