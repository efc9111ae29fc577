"""Build recipe of the HIP library ``libmkg.so`` (gfx950) and of the C++ shim.

Every source is compiled by ``hipcc`` in its own process (in parallel), then linked into
``libmkg.so`` next to this file.  Run ``python -m materkit_anisotropiclbm_trt_rbc_facet_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libmkg.so"
INCLUDE = HERE.parent / "include"

ARCH = os.environ.get("MKG_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from forming packed FP32 instructions.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

SOURCES = ["step_f32.hip", "step_f64.hip", "aux.hip", "reduce_static.hip", "adjoint.hip",
           "fadj.hip", "mpf.hip", "fvm.hip", "mp.hip",
           "api.cpp",
           "case.cpp"]

# Per-source flags.  The FP32 kernels divide and take square roots with the hardware
# reciprocal / square-root approximations (a few ulp) instead of the correctly rounded
# sequences, which cost about ten instructions each; the FP32 results stay within their
# tolerances against the FP64 oracle (1e-5 for fields, 1e-3 for gradients).  The flag leaves
# double arithmetic unchanged.  The finite-volume kernels lose 28 % (phase) and 18 % (solute)
# of their vector instructions with it and run 4 % faster at 256^3; the multi-phase and
# two-phase kernels, bound by HBM, measured no difference and keep the exact forms.
FAST_DIV = ["-fno-hip-fp32-correctly-rounded-divide-sqrt"]
EXTRA_FLAGS = {"step_f32.hip": FAST_DIV, "adjoint.hip": FAST_DIV, "fvm.hip": FAST_DIV}


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + "." + Path(src).suffix[1:] + ".o")
    cmd = [_hipcc(), *HIP_FLAGS, *EXTRA_FLAGS.get(src, []), "-c", str(CSRC / src), "-o", str(obj)]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


SHIM = HERE / "shim"
SHIM_LIB = HERE / "libmkshim.so"
MKRUN = HERE / "mkrun"


def build_shim() -> None:
    """The reference-API host layer (g++, host only) and the MaterKit-style driver."""
    cxx = shutil.which("g++") or "g++"
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{SHIM}"]
    tmp = SHIM_LIB.with_suffix(".so.tmp")
    cmds = [
        [cxx, *flags, "-shared", "-o", str(tmp), str(SHIM / "mkshim.cpp"), f"-L{HERE}", "-lmkg",
         "-Wl,-rpath,$ORIGIN"],
    ]
    for cmd in cmds:
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"shim build failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, SHIM_LIB)
    tmp = MKRUN.with_suffix(".tmp")
    cmd = [cxx, *flags, "-o", str(tmp), str(SHIM / "mkrun.cpp"), f"-L{HERE}", "-lmkshim", "-lmkg",
           "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, MKRUN)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source for gfx950 and link ``libmkg.so``; returns its path."""
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_shim()
    if verbose:
        print(f"built {LIB}, {SHIM_LIB}, {MKRUN}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
