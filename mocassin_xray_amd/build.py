"""Build recipes: libmocgpu.so (HIP, gfx950) and the Fortran drop-in shim.

The library is compiled in-tree so that the built .so travels with the
source tree.  Every HIP source is compiled with packed-FP32 generation off and
FP contraction off (the FP64 transport path then performs the same rounding
sequence as a plain IEEE CPU evaluation of the same formulas).
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libmocgpu.so"
FORTRAN = PKG / "fortran"

HIP_SOURCES = ["mocgpu.hip", "transport.hip", "opacity.hip", "balance.hip", "emission.hip", "dust.hip",
               "output.hip", "calib.hip"]
# host C++ of the same library: the native model set-up (mocmodel.h)
CPP_SOURCES = ["model.cpp", "model_data.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libmocgpu)")


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source of the package into libmocgpu.so."""
    srcs = [str(CSRC / s) for s in HIP_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    # the host sources go in as C++ (-x c++), the HIP sources after -x hip
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(ROOT / "include"), "-o", str(tmp),
           "-x", "c++", *[str(CSRC / s) for s in CPP_SOURCES], "-x", "hip", *srcs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stderr[-4000:])
    if verbose and res.stderr:
        print(res.stderr[-2000:])
    os.replace(tmp, LIB)
    return LIB


def flang() -> str | None:
    for cand in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang")):
        if cand and Path(cand).exists():
            return cand
    return None


def build_fortran_shim() -> Path | None:
    """Compile the ISO_C_BINDING module and its driver program (if flang exists)."""
    fc = flang()
    mod_src = FORTRAN / "mocgpu_mod.f90"
    drv_src = FORTRAN / "mocgpu_driver.f90"
    if fc is None or not mod_src.exists():
        return None
    out = FORTRAN / "mocgpu_driver"
    cmd = [fc, "-O2", "-module-dir", str(FORTRAN), "-o", str(out), str(mod_src), str(drv_src),
           f"-L{PKG}", "-lmocgpu", "-Wl,-rpath,$ORIGIN/.."]
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=str(FORTRAN))
    if res.returncode != 0:
        raise RuntimeError("flang failed:\n" + " ".join(cmd) + "\n" + res.stderr[-4000:])
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_fortran_shim())
