"""Build recipe of oracle/_ref/liboos_ref.so: the reference's `base` and `linalg` modules compiled from their sources
where they lie under /root/reference (nothing is copied into this repository), plus oracle/ref_driver.cpp, the C
functions the tests and the CPU-baseline leg of bench.py call. Test infrastructure only.

The flags follow the reference's own (cmake/OosKernels.cmake:15-40): everything at the x86-64 baseline with
-ffp-contract=off, the avx2 and avx512 kernel families with their target flags. The wasm128 family is left out
(src/linalg/CMakeLists.txt:4-9 builds it for WebAssembly only). Outputs go to oracle/_ref/ only.

    python oracle/build_ref.py            # incremental
    python oracle/build_ref.py --force    # rebuild every object
"""

import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = os.environ.get("OOS_REFERENCE_ROOT", "/root/reference")
OUT = os.path.join(HERE, "_ref")
OBJ = os.path.join(OUT, "obj")
LIB = os.path.join(OUT, "liboos_ref.so")
CXX = os.environ.get("OOS_REF_CXX", "/opt/rocm/llvm/bin/clang++")
BASE_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-march=x86-64", "-mtune=generic", "-ffp-contract=off",
              "-I" + os.path.join(REFERENCE, "include")]
FAMILY_FLAGS = {
    "avx2": ["-mavx2", "-mfma"],
    "avx512": ["-mavx512f", "-mavx512dq", "-mavx512vl", "-mavx2", "-mfma"],
}


def sources():
    """(source, extra flags) for every translation unit of base and linalg, as src/*/CMakeLists.txt globs them."""
    units = []
    for module in ("base", "linalg"):
        root = os.path.join(REFERENCE, "src", module)
        for folder, _, files in os.walk(root):
            for name in sorted(files):
                if not name.endswith(".cpp"):
                    continue
                path = os.path.join(folder, name)
                rel = os.path.relpath(path, os.path.join(REFERENCE, "src"))
                parts = rel.split(os.sep)
                extra = []
                if len(parts) > 3 and parts[1] == "kernels" and parts[2] in ("avx2", "avx512", "wasm128"):
                    if parts[2] == "wasm128":
                        continue
                    extra = FAMILY_FLAGS[parts[2]]
                units.append((path, extra))
    return units


def object_of(path):
    rel = os.path.relpath(path, os.path.join(REFERENCE, "src"))
    return os.path.join(OBJ, rel.replace(os.sep, "__")[:-4] + ".o")


def compile_one(job):
    path, extra, force = job
    obj = object_of(path)
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= os.path.getmtime(path):
        return obj, None
    cmd = [CXX] + BASE_FLAGS + extra + ["-c", path, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    return obj, (proc.stderr if proc.returncode != 0 else None)


def build(force=False, jobs=None):
    if not os.path.isdir(os.path.join(REFERENCE, "src", "linalg")):
        raise FileNotFoundError(f"the reference sources are not at {REFERENCE}")
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    units = sources()
    objects = []
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        for obj, error in pool.map(compile_one, [(p, e, force) for p, e in units]):
            if error:
                raise RuntimeError(f"compile failed for {obj}:\n{error}")
            objects.append(obj)
    driver = os.path.join(HERE, "ref_driver.cpp")
    cmd = [CXX] + BASE_FLAGS + ["-shared", "-o", LIB + ".tmp", driver] + objects + ["-lpthread"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("link of liboos_ref.so failed:\n" + proc.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    parser = argparse.ArgumentParser()
    parser.add_argument("--force", action="store_true")
    args = parser.parse_args()
    print(build(force=args.force))
    sys.exit(0)
