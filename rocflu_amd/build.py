"""Build recipes of the native libraries (in-tree, for gfx950).

- ``librocflu_amd.so``: HIP kernels + C-ABI (hipcc, gfx950, FMA contraction off so
  FP64 results round like the reference's gfortran build).
- ``librflu_mesh.so``: host region builder, grid and input file readers (g++).
- ``fortran/librflu_gpu_f.so`` + ``fortran/rflu_driver``: the ISO_C_BINDING module and
  a Fortran driver that calls the C-ABI in RungeKuttaMP's order (amdflang, when present);
  ``fortran/rflu_case_driver``: the same loop for a case read from the reference's input
  files;
  ``fortran/rflu_regions_driver``: the same loop over regions coupled through virtual cells.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}")


def build_hip():
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    src = os.path.join(HERE, "csrc", "capi.hip")
    out = os.path.join(HERE, "librocflu_amd.so")
    tmp = out + ".tmp"
    _run([hipcc] + HIPCC_FLAGS + ["-I" + INC, "-I" + os.path.join(HERE, "csrc"), src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_mesh():
    srcs = [os.path.join(HERE, "csrc", f) for f in ("mesh.cpp", "grid_io.cpp", "inputs.cpp", "sype.cpp")]
    out = os.path.join(HERE, "librflu_mesh.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wextra",
          "-I" + INC] + [s for s in srcs if os.path.exists(s)] + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_fortran():
    fc = shutil.which("amdflang") or os.path.join(ROCM, "bin", "amdflang")
    if not os.path.exists(fc):
        print("amdflang not found; skipping the Fortran binding", flush=True)
        return None
    fdir = os.path.join(HERE, "fortran")
    mod = os.path.join(fdir, "rflu_gpu.f90")
    drv = os.path.join(fdir, "rflu_driver.f90")
    if not (os.path.exists(mod) and os.path.exists(drv)):
        return None
    out = os.path.join(fdir, "rflu_driver")
    _run([fc, "-O2", "-J", fdir, mod, drv, "-L" + HERE, "-lrocflu_amd", "-lrflu_mesh",
          "-Wl,-rpath,$ORIGIN/..", "-o", out], cwd=fdir)
    # the host that runs a case from the reference's .inp / .bc / .hyb.asc files
    drv3 = os.path.join(fdir, "rflu_case_driver.f90")
    if os.path.exists(drv3):
        _run([fc, "-O2", "-J", fdir, mod, drv3, "-L" + HERE, "-lrocflu_amd", "-lrflu_mesh",
              "-Wl,-rpath,$ORIGIN/..", "-o", os.path.join(fdir, "rflu_case_driver")], cwd=fdir)
    # the host of several regions coupled through virtual cells
    drv2 = os.path.join(fdir, "rflu_regions_driver.f90")
    if os.path.exists(drv2):
        _run([fc, "-O2", "-J", fdir, mod, drv2, "-L" + HERE, "-lrocflu_amd",
              "-Wl,-rpath,$ORIGIN/..", "-o", os.path.join(fdir, "rflu_regions_driver")], cwd=fdir)
    return out


def build_checks():
    """GPU self-checks run by the tests (tests/native)."""
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    src = os.path.join(ROOT, "tests", "native", "inv_sqrt_check.hip")
    if not os.path.exists(src):
        return None
    out = src[:-len(".hip")]
    flags = [f for f in HIPCC_FLAGS if f not in ("-fPIC", "-shared")]
    _run([hipcc] + flags + ["-I" + INC, "-I" + os.path.join(HERE, "csrc"), src, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all(fortran: bool = True):
    build_mesh()
    build_hip()
    build_checks()
    if fortran:
        build_fortran()


if __name__ == "__main__":
    build_all(fortran="--no-fortran" not in sys.argv)
