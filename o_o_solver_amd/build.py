"""Builds the native libraries in-tree (they travel with the source tree; nothing is installed):

    o_o_solver_amd/liboosg.so   the product: HIP kernels for gfx950 + the host driver behind include/oosg.h + the
                                host analysis (ordering, symbolic factor)
    o_o_solver_amd/liboosm.so   the synthetic finite-element models the tests and the bench factorise
    oracle/liboracle.so         the CPU oracle (test infrastructure)
    oracle/_ref/liboos_ref.so   the compiled reference (test infrastructure; only where /root/reference exists)

    python -m o_o_solver_amd.build
"""

import concurrent.futures
import os
import subprocess
import sys

PACKAGE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PACKAGE)
CSRC = os.path.join(PACKAGE, "csrc")
BUILD = os.path.join(PACKAGE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions off for every HIP translation unit (hipcc flag; the host pass ignores it with a warning).
HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall"]
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall"]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + proc.stdout + proc.stderr)
    return proc


def _compile(job):
    source, obj, flags, compiler = job
    _run([compiler] + flags + ["-c", source, "-o", obj])
    return obj


def build_native(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = [
        (os.path.join(CSRC, "kernels.hip"), os.path.join(BUILD, "kernels.o"), HIP_FLAGS, HIPCC),
        (os.path.join(CSRC, "selector.hip"), os.path.join(BUILD, "selector.o"), HIP_FLAGS, HIPCC),
        (os.path.join(CSRC, "pivoting.hip"), os.path.join(BUILD, "pivoting.o"), HIP_FLAGS, HIPCC),
        (os.path.join(CSRC, "scratch.hip"), os.path.join(BUILD, "scratch.o"), HIP_FLAGS, HIPCC),
        (os.path.join(CSRC, "oosg.cpp"), os.path.join(BUILD, "oosg.o"), HOST_FLAGS + ["-D__HIP_PLATFORM_AMD__"],
         HIPCC),
        (os.path.join(CSRC, "analysis.cpp"), os.path.join(BUILD, "analysis.o"), HOST_FLAGS, "g++"),
        (os.path.join(CSRC, "models.cpp"), os.path.join(BUILD, "models.o"), HOST_FLAGS, "g++"),
    ]
    with concurrent.futures.ThreadPoolExecutor(len(jobs)) as pool:
        objects = list(pool.map(_compile, jobs))
    lib = os.path.join(PACKAGE, "liboosg.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", lib + ".tmp", *objects[:6]])
    os.replace(lib + ".tmp", lib)
    models = os.path.join(PACKAGE, "liboosm.so")
    _run(["g++", "-shared", "-fPIC", "-o", models + ".tmp", objects[6]])
    os.replace(models + ".tmp", models)
    if verbose:
        print(lib)
        print(models)
    return lib


def build_oracle(verbose=False):
    oracle = os.path.join(ROOT, "oracle")
    lib = os.path.join(oracle, "liboracle.so")
    _run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-o", lib + ".tmp",
          os.path.join(oracle, "oracle.c"), "-lm"])
    os.replace(lib + ".tmp", lib)
    if verbose:
        print(lib)
    reference = os.environ.get("OOS_REFERENCE_ROOT", "/root/reference")
    if os.path.isdir(os.path.join(reference, "src", "linalg")):
        sys.path.insert(0, oracle)
        try:
            import build_ref
            out = build_ref.build()
            if verbose:
                print(out)
        finally:
            sys.path.pop(0)
    return lib


def build_all(verbose=False):
    build_native(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
