"""The C ABI: liboosg.so loads and exports every function include/oosg.h declares (no GPU call)."""

import ctypes
import os
import subprocess

from o_o_solver_amd import _native


def test_library_exports_every_declared_function():
    declared = _native.header_functions()
    assert "oosg_factorise" in declared and "oosg_solve" in declared and len(declared) >= 20
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [name for name in declared if not hasattr(lib, name)]
    assert missing == []


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for name in _native.header_functions():
        assert name in exported, name


def test_header_compiles_as_c():
    header = os.path.join(os.path.dirname(_native.LIB_PATH), "..", "include", "oosg.h")
    src = f'#include "{os.path.abspath(header)}"\nint main(void) {{ return (int)OOSG_OK; }}\n'
    proc = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-x", "c", "-fsyntax-only", "-"], input=src,
                          capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr


def test_last_error_is_a_string():
    assert isinstance(_native.lib().oosg_last_error(), bytes)
