"""Build the HIP step library in-tree: simson_amd/_lib/libsimson_gpu.so.

    python -m simson_amd.build            # gfx950, all sources

Every source under simson_amd/csrc is compiled for gfx950 with hipcc and
linked into one shared library exporting the C ABI of include/simson_gpu.h.
Packed-FP32 instructions are disabled (-packed-fp32-ops): some of their
forms have given wrong results on the target machines without an error.

The Fortran host (simson_amd/fortran: the ISO_C_BINDING module and the
blagpu driver that reads bla.i and .u files) is compiled with amdflang into
simson_amd/_lib/blagpu, linked to the step library by rpath.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUTDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUTDIR, "libsimson_gpu.so")
ARCH = os.environ.get("SIMSON_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-result",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OUTDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        # keep warnings visible but ignore the host pass not knowing the
        # GPU-only target feature
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l]
        if any(l.strip() for l in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


CALIB_LIB = os.path.join(OUTDIR, "libsimson_calib.so")


def _link(objs, out):
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)


FORTRAN = os.path.join(HERE, "fortran")
BLAGPU = os.path.join(OUTDIR, "blagpu")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")


def build_fortran() -> str:
    """Compile the Fortran host against the step library (rpath $ORIGIN)."""
    mod = os.path.join(FORTRAN, "simson_gpu_mod.f90")
    prog = os.path.join(FORTRAN, "blagpu.f90")
    tmp = BLAGPU + ".tmp"
    cmd = [FC, "-O2", "-J", OUTDIR, mod, prog, "-o", tmp, "-L", OUTDIR, "-lsimson_gpu",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, BLAGPU)
    return BLAGPU


def build(verbose: bool = True) -> str:
    """Compile the step library, the copy-calibration library and the
    Fortran host."""
    os.makedirs(OUTDIR, exist_ok=True)
    srcs = sources()
    calib = [os.path.join(CSRC, "calib", "sg_copy.hip")]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs) + 1)) as ex:
        objs = list(ex.map(_compile, srcs + calib))
    _link(objs[: len(srcs)], LIB)
    _link(objs[len(srcs):], CALIB_LIB)
    build_fortran()
    if verbose:
        print(f"built {LIB}, {CALIB_LIB} and {BLAGPU}")
    return LIB


if __name__ == "__main__":
    build()
