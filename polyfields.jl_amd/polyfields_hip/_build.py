"""Compile the HIP library ``libpolyfields_hip.so`` for gfx950 (MI355X).

Usage: ``python -m polyfields_hip._build`` (from ``polyfields.jl_amd/``) or
``polyfields_hip._build.build()``.  Objects go to ``polyfields.jl_amd/_build/``; the
shared library is written next to this file, where the ctypes loader looks for it.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, "csrc")
OBJ = os.path.join(PKG_ROOT, "_build")
LIB = os.path.join(HERE, "libpolyfields_hip.so")

SOURCES = ["pf_fft.hip", "pf_fft_fixed.hip", "pf_kernels.hip", "pf_batch_kernels.hip", "pf_engine.hip", "pf_adjoint.hip", "pf_noalloc.hip", "pf_batch.cpp", "pf_system.cpp",
           "pf_grad.cpp", "pf_tune.cpp", "pf_cell.cpp", "pf_julia_random.cpp", "pf_api.cpp"]

ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")
# Packed FP32 instructions are disabled for all device code (they are generated by the
# compiler from ordinary float code otherwise).
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found; install ROCm or put hipcc on PATH")
    return h


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    p = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                     if "unknown target feature" not in l and "packed-fp32-ops" not in l
                     and l.strip() and not l.startswith("1 warning generated"))
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{msgs}")
    return out, msgs


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for _, m in results:
            if m:
                print(m)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
