"""Builds the gfx950 front-end library (jixel_amd/libjxg.so) in-tree.

hipcc compiles the kernels for gfx950 only; the host-side parameter code is
plain C++.  Floating-point contraction is disabled everywhere so that every
a*b+c is rounded exactly as the reference rounds it, and packed-FP32
instructions are disabled for the device code.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libjxg.so")
STATIC = os.path.join(HERE, "libjxg.a")  # the same objects, for a static link (link -lamdhip64 too)
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
          "-I", os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=gfx950", "-fno-gpu-flush-denormals-to-zero",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


HIP_SOURCES = ["jxg_kernels.hip", "jxg_filters.hip", "jxg_cfl.hip", "jxg_large.hip",
               "jxg_epf.hip", "jxg_alpha.hip", "jxg_chroma.hip",
               "jxg_batch.hip", "jxg_tune.hip", "jxg_retile.hip", "jxg_refine.hip"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


HOST_SOURCES = ["jxg_params.cpp", "jxg_tables.cpp"]


def build(verbose=True, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    cmds, objs = [], []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        cmds.append([CXX] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmds.append([HIPCC] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the translation units are independent: compile them side by side
    jobs = jobs or min(len(cmds), max(1, (os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    tmp = STATIC + ".tmp"
    if os.path.exists(tmp):
        os.remove(tmp)
    run(["ar", "rcs", tmp] + objs)
    os.replace(tmp, STATIC)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
