"""Build the native libraries in-tree (amrex_amd/_lib) for gfx950.

* libmgx.so          HIP kernels + C++ MLMG + C ABI (hipcc, --offload-arch=gfx950)
* libmgx_workload.so seeded input generator (g++ -fopenmp)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "_lib")
OBJ = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are switched off for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
SOURCES = ["mgx_kernels.hip", "mgx_smoother.hip", "mgx_sweep2.hip", "mgx_fused.hip", "mgx_march.hip", "mgx_amr.hip", "mgx_krylov.hip",
           "mgx_io.hip", "mgx_mfreduce.hip", "mgx_probe.hip", "mlmg_host.cpp", "mlmg_solve.cpp", "gmres_mlmg.cpp", "plotfile_io.cpp", "fi_capi.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def build_workload():
    os.makedirs(LIB, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o",
          os.path.join(LIB, "libmgx_workload.so"), os.path.join(CSRC, "workload.cpp")])


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, cmds))
    tmp = os.path.join(LIB, "libmgx.so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, os.path.join(LIB, "libmgx.so"))


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_workload()
    build_hip()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", os.listdir(LIB))
