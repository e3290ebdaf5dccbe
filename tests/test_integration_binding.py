"""The reference-side binding shown in INTEGRATION.md compiles against the reference's own headers and include/oosg.h
(a syntax-only compile: the binding is the C++ a maintainer adds to the reference, so its identifiers, constructors
and fields must be the reference's). Skipped where the reference's headers are not present."""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE_INCLUDE = os.path.join(os.environ.get("OOS_REFERENCE_ROOT", "/root/reference"), "include")


def _binding():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    assert blocks, "INTEGRATION.md has no C++ binding"
    return max(blocks, key=len)


def _compiler():
    for name in ("/opt/rocm/llvm/bin/clang++", "clang++"):
        path = shutil.which(name) or (name if os.path.exists(name) else None)
        if path:
            return path
    return None


def test_binding_compiles_against_the_reference_headers(tmp_path):
    if not os.path.isdir(os.path.join(REFERENCE_INCLUDE, "oos")):
        pytest.skip("the reference's headers are not here")
    compiler = _compiler()
    if compiler is None:
        pytest.skip("no clang++")
    source = tmp_path / "GpuFactor.cpp"
    source.write_text(_binding())
    proc = subprocess.run([compiler, "-std=c++20", "-fsyntax-only", "-Wall", "-Werror", "-I", REFERENCE_INCLUDE,
                           "-I", os.path.join(ROOT, "include"), str(source)], capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-4000:]


def test_binding_calls_only_declared_entry_points():
    from o_o_solver_amd import _native
    declared = set(_native.header_functions())
    called = set(re.findall(r"\b(oosg_[a-z_0-9]+)\s*\(", _binding()))
    assert called and called <= declared, called - declared
