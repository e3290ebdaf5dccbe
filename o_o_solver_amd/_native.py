"""ctypes bindings of include/oosg.h (liboosg.so) and of the model generator (liboosm.so).

The product path has no fallback: when the native library is missing, the import fails with the build command.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboosg.so")
MODELS_PATH = os.path.join(_HERE, "liboosm.so")

c_void_p = ctypes.c_void_p
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_int = ctypes.c_int


class SymbolicDesc(ctypes.Structure):
    _fields_ = [("n", c_int32), ("supernode_count", c_int32), ("first_column", c_void_p),
                ("column_count", c_void_p), ("parent", c_void_p), ("row_start", c_void_p), ("rows", c_void_p)]


class BlockMatrixDesc(ctypes.Structure):
    _fields_ = [("node_count", c_int32), ("first_dof", c_void_p), ("column_start", c_void_p),
                ("block_rows", c_void_p), ("block_value_offsets", c_void_p), ("values", c_void_p)]


class OutOfCoreTrafficC(ctypes.Structure):
    _fields_ = [("spooled_bytes", ctypes.c_uint64), ("written_bytes", ctypes.c_uint64),
                ("read_bytes", ctypes.c_uint64), ("read_requests", ctypes.c_uint64)]


class CscDesc(ctypes.Structure):
    _fields_ = [("n", c_int32), ("offsets", c_void_p), ("rows", c_void_p), ("values", c_void_p)]


class FactorOptionsC(ctypes.Structure):
    _fields_ = [("pivoting", c_int32), ("precision", c_int32), ("max_ratio", c_double), ("bailout", c_int32),
                ("reference_diagonal", c_void_p), ("fp32_levels", c_int32)]


class FactorInfoC(ctypes.Structure):
    _fields_ = [("positive", c_int64), ("negative", c_int64), ("zero", c_int64), ("delayed_pivots", c_int64),
                ("two_by_two_pivots", c_int64), ("issue_count", c_int64), ("first_issue_equation", c_int32),
                ("first_issue_negative", c_int32), ("largest_ratio", c_double), ("largest_ratio_equation", c_int32),
                ("precision", c_int32), ("factor_seconds", c_double), ("flops", c_double), ("fp32_levels", c_int32)]


class SolveReportC(ctypes.Structure):
    _fields_ = [("steps", c_int32), ("converged", c_int32), ("backward_error", c_double), ("seconds", c_double)]


class PivotIssueC(ctypes.Structure):
    _fields_ = [("equation", c_int32), ("kind", c_int32), ("matrix_diagonal", c_double),
                ("factor_diagonal", c_double), ("ratio", c_double)]


class MemoryShortfallC(ctypes.Structure):
    _fields_ = [("requested_bytes", ctypes.c_uint64), ("available_bytes", ctypes.c_uint64),
                ("limit_bytes", ctypes.c_uint64)]


CANCEL_FN = ctypes.CFUNCTYPE(c_int32, c_void_p)


_SIGNATURES = {
    "oosg_last_error": (ctypes.c_char_p, []),
    "oosg_last_error_equations": (c_int, [c_void_p, c_int64, ctypes.POINTER(c_int64)]),
    "oosg_last_error_shortfall": (c_int, [ctypes.POINTER(MemoryShortfallC)]),
    "oosg_estimate_memory_bytes": (c_int, [ctypes.POINTER(SymbolicDesc), c_int32, ctypes.POINTER(ctypes.c_uint64)]),
    "oosg_plan_memory": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "oosg_plan_set_stream": (c_int, [c_void_p, c_void_p]),
    "oosg_plan_set_cancellation": (c_int, [c_void_p, CANCEL_FN, c_void_p]),
    "oosg_factor_issues": (c_int, [c_void_p, c_void_p, c_int64, ctypes.POINTER(c_int64)]),
    "oosg_plan_create": (c_int, [ctypes.POINTER(SymbolicDesc), c_int32, ctypes.c_uint64, ctypes.POINTER(c_void_p)]),
    "oosg_plan_create_partial": (c_int, [ctypes.POINTER(SymbolicDesc), c_int32, c_int32, ctypes.c_uint64,
                                         ctypes.POINTER(c_void_p)]),
    "oosg_partial_schur_complement": (c_int, [c_void_p, c_void_p, c_int64]),
    "oosg_plan_create_out_of_core": (c_int, [ctypes.POINTER(SymbolicDesc), c_int32, ctypes.c_uint64,
                                             ctypes.POINTER(c_void_p)]),
    "oosg_factor_out_of_core_traffic": (c_int, [c_void_p, ctypes.POINTER(OutOfCoreTrafficC)]),
    "oosg_partial_condense_loads": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_int64]),
    "oosg_partial_recover": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_int64]),
    "oosg_csc_checksum": (c_int, [ctypes.POINTER(CscDesc), ctypes.POINTER(ctypes.c_uint64)]),
    "oosg_csc_from_blocks_device": (c_int, [c_int32, ctypes.POINTER(BlockMatrixDesc), c_void_p, c_int64, c_int32,
                                            c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_int64), c_void_p]),
    "oosg_matrix_create_device": (c_int, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                                          ctypes.POINTER(c_void_p)]),
    "oosg_matrix_set_values_device": (c_int, [c_void_p, c_void_p]),
    "oosg_plan_destroy": (None, [c_void_p]),
    "oosg_plan_stats": (c_int, [c_void_p, c_void_p]),
    "oosg_plan_synchronize": (c_int, [c_void_p]),
    "oosg_plan_use_graphs": (c_int, [c_void_p, c_int32]),
    "oosg_plan_tile_policy": (c_int, [c_void_p, c_int32]),
    "oosg_plan_streams": (c_int, [c_void_p, c_int32, c_int32]),
    "oosg_plan_launch_mode": (c_int, [c_void_p, c_int32, c_int32]),
    "oosg_plan_get_launch_mode": (c_int, [c_void_p, c_int32, c_void_p]),
    "oosg_plan_launch_times": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "oosg_plan_autotune": (c_int, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    "oosg_plan_profile": (c_int, [c_void_p, c_int32]),
    "oosg_plan_profile_read": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "oosg_matrix_create": (c_int, [c_void_p, ctypes.POINTER(CscDesc), ctypes.POINTER(c_void_p)]),
    "oosg_matrix_set_values": (c_int, [c_void_p, c_void_p]),
    "oosg_matrix_set_combination": (c_int, [c_void_p, c_double, c_void_p, c_double, c_void_p, c_void_p]),
    "oosg_matrix_destroy": (None, [c_void_p]),
    "oosg_factorise": (c_int, [c_void_p, c_void_p, ctypes.POINTER(FactorOptionsC), ctypes.POINTER(c_void_p)]),
    "oosg_factor_destroy": (None, [c_void_p]),
    "oosg_factor_info_get": (c_int, [c_void_p, ctypes.POINTER(FactorInfoC)]),
    "oosg_factor_panel": (c_int, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    "oosg_factor_diagonal": (c_int, [c_void_p, c_void_p]),
    "oosg_factor_panel_shape": (c_int, [c_void_p, c_int32, c_void_p, c_void_p]),
    "oosg_factor_subdiagonal": (c_int, [c_void_p, c_int32, c_void_p]),
    "oosg_solve": (c_int, [c_void_p, c_void_p, c_int64, c_int64]),
    "oosg_solve_device": (c_int, [c_void_p, c_void_p, c_int64, c_int64]),
    "oosg_solve_refined_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_double,
                                          c_int32, ctypes.POINTER(SolveReportC)]),
    "oosg_solve_refined_device_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64,
                                                  c_double, c_int32, ctypes.POINTER(SolveReportC)]),
    "oosg_factorise_solve_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64,
                                            c_double, c_int32, c_void_p, ctypes.POINTER(SolveReportC)]),
    "oosg_matrix_multiply_device": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64]),
    "oosg_debug_mfma_probe": (c_int, [c_void_p] * 6),
    "oosg_selector_create": (c_int, [c_int32, ctypes.c_uint64, ctypes.POINTER(c_void_p)]),
    "oosg_selector_destroy": (None, [c_void_p]),
    "oosg_selector_parameters": (c_int, [c_void_p, c_void_p]),
    "oosg_selector_set_parameters": (c_int, [c_void_p, c_void_p]),
    "oosg_selector_cut_features": (c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "oosg_selector_predict": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "oosg_selector_gradient": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "oosg_selector_train": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int32, ctypes.c_float, c_void_p]),
    "oosa_order_dissection": (c_int, [c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "oosa_analyse": (c_void_p, [c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int]),
    "oosa_free": (None, [c_void_p]),
    "oosa_sizes": (None, [c_void_p, c_void_p, c_void_p]),
    "oosa_export": (None, [c_void_p] * 7),
    "oosa_memory": (None, [c_void_p, c_void_p]),
    "oosa_last_error": (ctypes.c_char_p, []),
}

_MODEL_SIGNATURES = {
    "oosm_plate": (c_void_p, [c_int32, c_int32, c_double, c_double, c_double, ctypes.c_uint64, c_double, c_double,
                              c_int, c_double, c_double, c_double]),
    "oosm_hex_block": (c_void_p, [c_int32, c_int32, c_int32, c_double, c_double, c_double, ctypes.c_uint64,
                                  c_double, c_double, c_double]),
    "oosm_free": (None, [c_void_p]),
    "oosm_sizes": (None, [c_void_p, c_void_p]),
    "oosm_graph": (None, [c_void_p, c_void_p, c_void_p]),
    "oosm_coords": (None, [c_void_p, c_void_p]),
    "oosm_csc": (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "oosm_csc_geometric": (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "oosm_cylinder": (c_void_p, [c_int32, c_int32, c_double, c_double, c_double, c_double, c_double, c_int32,
                                 c_double, c_double]),
    "oosm_loads": (None, [c_void_p, c_void_p, c_int, c_int32, ctypes.c_uint64, c_void_p]),
}

_lib = None
_models = None


def _bind(path, signatures):
    lib = ctypes.CDLL(path)
    for name, (restype, argtypes) in signatures.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def lib():
    """liboosg.so, loaded once. Raises ImportError (with the build command) when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m o_o_solver_amd.build` (hipcc, gfx950)")
        _lib = _bind(LIB_PATH, _SIGNATURES)
    return _lib


def models():
    global _models
    if _models is None:
        if not os.path.exists(MODELS_PATH):
            raise ImportError(f"{MODELS_PATH} is not built: run `python -m o_o_solver_amd.build`")
        _models = _bind(MODELS_PATH, _MODEL_SIGNATURES)
    return _models


def header_functions(header_path=None):
    """The function names include/oosg.h declares (for the export test)."""
    import re
    header_path = header_path or os.path.join(os.path.dirname(_HERE), "include", "oosg.h")
    text = open(header_path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oos[ag]_[a-z_0-9]+)\s*\(", text)))
