"""In-tree build of the native libraries (no JIT cache, no site-packages).

  fasteig_amd/lib/libfe_hip.so       gfx950 kernels + the C ABI of include/fe_hip.h
  fasteig_amd/lib/libfasteig_hip.so  C++ host layer (fasteig_problem / options /
                                     solver API, fasteig::hip backend, solvers,
                                     serving loop) over libfe_hip.so
  fasteig_amd/lib/runfasteig_hip     input-deck driver executable

Run ``python -m fasteig_amd.build`` (or __graft_entry__.build()).
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(CSRC, "host")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-parameter", "-fopenmp",
              "-I", os.path.join(ROOT, "include"), "-I", HOST]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def _compile_hip(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    return obj


def _compile_host(src):
    obj = os.path.join(OBJ, "host_" + os.path.basename(src) + ".o")
    _run(["g++", *HOST_FLAGS, "-c", src, "-o", obj])
    return obj


def build(verbose: bool = True) -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    hip_srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    host_srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    driver = os.path.join(CSRC, "runfasteig.cpp")
    workers = min(8, max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(workers) as ex:
        hip_objs = list(ex.map(_compile_hip, hip_srcs))
        host_objs = list(ex.map(_compile_host, host_srcs))
    fe_so = os.path.join(LIB, "libfe_hip.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", fe_so + ".tmp", *hip_objs])
    os.replace(fe_so + ".tmp", fe_so)
    if host_objs:
        host_so = os.path.join(LIB, "libfasteig_hip.so")
        _run(["g++", "-shared", "-fPIC", "-fopenmp", "-o", host_so + ".tmp", *host_objs, "-L", LIB, "-lfe_hip",
              "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
        os.replace(host_so + ".tmp", host_so)
        if os.path.exists(driver):
            exe = os.path.join(LIB, "runfasteig_hip")
            _run(["g++", *HOST_FLAGS, "-o", exe + ".tmp", driver, "-L", LIB, "-lfasteig_hip", "-lfe_hip",
                  "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
            os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {fe_so} ({len(hip_srcs)} HIP sources, {len(host_srcs)} host sources)")


if __name__ == "__main__":
    build()
