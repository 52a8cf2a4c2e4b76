"""Builds libpangu_hip.so in-tree for gfx950.

Host files that do physics arithmetic (geometry tables, initial data) are
compiled with g++ -ffp-contract=off so their FP64 results are reproducible;
HIP sources are compiled with hipcc for gfx950 with contraction off as well
(-ffp-contract=off) and packed-FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpangu_hip.so")
PHDF_LIB = os.path.join(HERE, "libpangu_phdf.so")
# HDF5 (C library) for the Parthenon-style output writer
HDF5_ROOT = os.environ.get("PANGU_HDF5_ROOT", "/opt/conda")
ARCH = os.environ.get("PANGU_OFFLOAD_ARCH", "gfx950")

HOST_SRC = ["geometry.cpp", "problems.cpp", "refine_host.cpp"]
HIP_SRC = ["kernels.hip", "api.hip", "sim.hip", "mlp.hip", "diag.hip", "refine.hip"]

GXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
             "-Wall", "-Wno-unused-variable"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r


def build(jobs=8, verbose=False):
    """Compile every source and link the shared library (always rebuilds)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    os.makedirs(BUILD, exist_ok=True)
    procs = []
    objs = []
    for src in HOST_SRC:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        procs.append(["g++", *GXX_FLAGS, "-c", src, "-o", obj])
    for src in HIP_SRC:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        procs.append([hipcc, *HIP_FLAGS, "-c", src, "-o", obj])
    running = []
    for cmd in procs:
        if verbose:
            print(" ".join(cmd))
        running.append((cmd, subprocess.Popen(cmd, cwd=CSRC, stdout=subprocess.PIPE,
                                               stderr=subprocess.PIPE, text=True)))
        while len([p for _, p in running if p.poll() is None]) >= jobs:
            running[0][1].wait()
    failed = False
    for cmd, p in running:
        out, err = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(" ".join(cmd) + "\n" + out + err)
    if failed:
        raise RuntimeError("compilation failed")
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lgomp"])
    os.replace(tmp, LIB)
    build_phdf()
    return LIB


def build_phdf():
    """libpangu_phdf.so (host only, links HDF5); skipped when the HDF5
    headers are absent, in which case pangu_amd.phdf raises on use."""
    inc = os.path.join(HDF5_ROOT, "include")
    libdir = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        sys.stderr.write(f"pangu_amd.build: no hdf5.h under {inc}; libpangu_phdf.so not built\n")
        return None
    tmp = PHDF_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{inc}", "phdf.cpp", "-o", tmp,
          f"-L{libdir}", "-lhdf5", f"-Wl,-rpath,{libdir}"])
    os.replace(tmp, PHDF_LIB)
    return PHDF_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
