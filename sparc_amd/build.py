"""In-tree build of the HIP engine (gfx950) and, for the tests, the CPU oracle.

``python -m sparc_amd.build`` compiles ``sparc_amd/libsparc_gpu.so`` with hipcc;
``__graft_entry__.build()`` calls :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed FP32 instructions are kept out of the generated code
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = [os.path.join(HERE, "csrc", "sparc_gpu.hip")]
DEPS = SOURCES + [os.path.join(HERE, "csrc", f)
                  for f in ("kernels.h", "fft_lds.h", "fft_reg.h", "fft_r16.h", "zsolve.h",
                            "lowword.h")] + [
    os.path.join(ROOT, "include", "sparc_gpu.h")]


def build_engine(verbose: bool = False) -> str:
    out = os.path.join(HERE, "libsparc_gpu.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *HIPFLAGS, "-o", tmp, *SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libsparc_gpu.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, out)
    return out


def build_oracle() -> None:
    """Builds oracle/liboracle.so and, when the reference tree is present,
    oracle/_ref (test infrastructure only)."""
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")


FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
FORTRAN_EXE = os.path.join(HERE, "fortran", "sparc_timestepping")


def build_fortran() -> str | None:
    """The Fortran host (include/sparc_gpu_mod.f90 + fortran/timestepping.f90)
    linked against libsparc_gpu.so with an $ORIGIN rpath; skipped when no
    Fortran compiler is present."""
    if not os.path.exists(FC):
        return None
    moddir = os.path.join(HERE, "fortran", "_mod")
    os.makedirs(moddir, exist_ok=True)
    cmd = [FC, "-O2", "-J", moddir, "-o", FORTRAN_EXE,
           os.path.join(ROOT, "include", "sparc_gpu_mod.f90"),
           os.path.join(HERE, "fortran", "timestepping.f90"),
           "-L" + HERE, "-lsparc_gpu", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("amdflang failed building the Fortran host")
    return FORTRAN_EXE


def build_all() -> None:
    build_engine()
    build_fortran()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(HERE, "libsparc_gpu.so"))
