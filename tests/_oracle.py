"""ctypes access to the checkers: oracle/liboracle.so (the C restatement) and oracle/_ref/liboos_ref.so (the compiled
reference, present where it was built). Test infrastructure only."""

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "liboos_ref.so")

vp = ctypes.c_void_p
i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double


def P(a):
    return None if a is None else a.ctypes.data_as(vp)


class OrcSymbolic(ctypes.Structure):
    _fields_ = [("n", i32), ("supernode_count", i32), ("first_column", vp), ("column_count", vp), ("parent", vp),
                ("row_start", vp), ("rows", vp)]


class OrcResult(ctypes.Structure):
    _fields_ = [("status", i32), ("zero_pivot_count", i32), ("first_zero_pivot", i32), ("first_issue", i32),
                ("first_issue_negative", i32), ("largest_ratio_equation", i32), ("issue_count", i64),
                ("inertia", i64 * 3), ("largest_ratio", dbl)]


_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        lib = ctypes.CDLL(ORACLE_PATH)
        lib.orc_ldlt_nopivot.restype = i64
        lib.orc_ldlt_nopivot.argtypes = [vp, i64, i64, i64]
        lib.orc_front_nopivot.restype = i64
        lib.orc_front_nopivot.argtypes = [vp, i64, vp, i64, i64, i64]
        lib.orc_bunch_kaufman.restype = i64
        lib.orc_bunch_kaufman.argtypes = [vp, i64, i64, i64, vp, vp]
        lib.orc_inertia.restype = None
        lib.orc_inertia.argtypes = [vp, vp, i64, vp]
        lib.orc_spmv.restype = None
        lib.orc_spmv.argtypes = [i32, vp, vp, vp, vp, vp]
        lib.orc_multifrontal_pd.restype = None
        lib.orc_multifrontal_pd.argtypes = [ctypes.POINTER(OrcSymbolic), vp, vp, vp, dbl, vp, vp,
                                            ctypes.POINTER(OrcResult)]
        lib.orc_solve.restype = None
        lib.orc_solve.argtypes = [ctypes.POINTER(OrcSymbolic), vp, vp, i64, i64]
        lib.orc_multifrontal_partial.restype = None
        lib.orc_multifrontal_partial.argtypes = [ctypes.POINTER(OrcSymbolic), vp, vp, vp, dbl, vp, vp, vp,
                                                 ctypes.POINTER(OrcResult)]
        lib.orc_partial_condense.restype = None
        lib.orc_partial_condense.argtypes = [ctypes.POINTER(OrcSymbolic), vp, vp, i64, i64]
        lib.orc_csc_from_blocks.restype = i64
        lib.orc_csc_from_blocks.argtypes = [i32, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp]
        lib.orc_partial_recover.restype = None
        lib.orc_partial_recover.argtypes = [ctypes.POINTER(OrcSymbolic), vp, vp, i64, i64, vp, i64]
        _orc = lib
    return _orc


def reference():
    """The compiled reference, or None where it was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        lib = ctypes.CDLL(REF_PATH)
        lib.ref_analyse.restype = vp
        lib.ref_analyse.argtypes = [i32, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp,
                                    ctypes.c_char_p, ctypes.c_int]
        lib.ref_sym_free.argtypes = [vp]
        lib.ref_sym_sizes.argtypes = [vp, vp, vp]
        lib.ref_sym_export.argtypes = [vp] * 7
        lib.ref_factor_solve.restype = ctypes.c_int
        lib.ref_factor_solve.argtypes = [vp, i32, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dbl,
                                         ctypes.c_int, vp, i64, vp, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]
        lib.ref_spmv.restype = ctypes.c_int
        lib.ref_spmv.argtypes = [i32, vp, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]
        lib.ref_dense_front_nopivot.restype = i64
        lib.ref_dense_front_nopivot.argtypes = [vp, i64, vp, i64, i64, i64]
        lib.ref_dense_bunch_kaufman.restype = i64
        lib.ref_dense_bunch_kaufman.argtypes = [vp, i64, i64, i64, vp, vp]
        lib.ref_hardware_threads.restype = ctypes.c_int
        lib.ref_combine.restype = ctypes.c_int
        lib.ref_combine.argtypes = [i32, vp, vp, dbl, vp, dbl, vp, vp, ctypes.c_char_p, ctypes.c_int]
        lib.ref_scale_symmetric.restype = ctypes.c_int
        lib.ref_scale_symmetric.argtypes = [i32, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]
        lib.ref_pencil_scale_of.restype = ctypes.c_int
        lib.ref_pencil_scale_of.argtypes = [i32, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]
        lib.ref_analyse_trailing.restype = vp
        lib.ref_analyse_trailing.argtypes = [i32, vp, vp, vp, vp, i32, ctypes.c_int, vp, ctypes.c_char_p,
                                             ctypes.c_int]
        lib.ref_pattern_export.restype = ctypes.c_int
        lib.ref_pattern_export.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]
        lib.ref_selection.restype = ctypes.c_int
        lib.ref_selection.argtypes = [i32, vp, vp, i64, vp, i32, vp, vp, ctypes.c_char_p, ctypes.c_int]
        lib.ref_from_blocks.restype = ctypes.c_int
        lib.ref_from_blocks.argtypes = [i32, vp, vp, vp, vp, vp, i64, vp, i32, ctypes.c_int, vp, vp, vp, vp,
                                        ctypes.c_char_p, ctypes.c_int]
        lib.ref_partial.restype = ctypes.c_int
        lib.ref_partial.argtypes = [vp, i32, vp, vp, vp, ctypes.c_int, dbl, vp, vp, vp, vp, vp, i64, vp, vp, vp,
                                    ctypes.c_char_p, ctypes.c_int]
        _ref = lib
    return _ref


def orc_symbolic(sym):
    """An OrcSymbolic over a SymbolicFactor's arrays (keep `sym` alive while it is used)."""
    return OrcSymbolic(sym.n, sym.supernode_count, P(sym.first_column), P(sym.column_count), P(sym.parent),
                       P(sym.row_start), P(sym.rows_all))


def multifrontal(sym, matrix, max_ratio=1e7):
    """(panels, D per equation, OrcResult) of the oracle's positive definite multifrontal factorisation."""
    lib = oracle()
    panels = np.zeros(sym.panel_value_count)
    diag = np.zeros(sym.n)
    res = OrcResult()
    s = orc_symbolic(sym)
    lib.orc_multifrontal_pd(ctypes.byref(s), P(matrix.offsets), P(matrix.rows), P(matrix.values), max_ratio,
                            P(panels), P(diag), ctypes.byref(res))
    return panels, diag, res


def oracle_solve(sym, panels, b):
    x = np.array(b, dtype=np.float64, order="F", copy=True).reshape(sym.n, -1, order="F")
    s = orc_symbolic(sym)
    oracle().orc_solve(ctypes.byref(s), P(panels), P(x), sym.n, x.shape[1])
    return x


def oracle_spmv(matrix, x):
    y = np.zeros(matrix.n)
    x = np.ascontiguousarray(x, dtype=np.float64)
    oracle().orc_spmv(matrix.n, P(matrix.offsets), P(matrix.rows), P(matrix.values), P(x), P(y))
    return y


def ref_symbolic(graph_offsets, adjacency, weights, coords=None, order=None, method=4, amalgamate=1, threads=1):
    """The reference's SymbolicFactor::analyse as (handle, sizes, flops, arrays dict); free with ref_sym_free."""
    lib = reference()
    n = int(len(graph_offsets) - 1)
    status = ctypes.c_int()
    err = ctypes.create_string_buffer(1024)
    h = lib.ref_analyse(n, P(graph_offsets), P(adjacency), P(weights), P(coords), P(order), method, amalgamate,
                        threads, ctypes.byref(status), err, 1024)
    if status.value != 0:
        raise RuntimeError(err.value.decode())
    sizes = np.zeros(7, np.int64)
    flops = dbl()
    lib.ref_sym_sizes(h, P(sizes), ctypes.byref(flops))
    S = int(sizes[2])
    arrays = dict(node_order=np.zeros(n, np.int32), first_column=np.zeros(S, np.int32),
                  column_count=np.zeros(S, np.int32), parent=np.zeros(S, np.int32),
                  row_start=np.zeros(S + 1, np.int64), rows=np.zeros(int(sizes[3]), np.int32))
    lib.ref_sym_export(h, *[P(arrays[k]) for k in ["node_order", "first_column", "column_count", "parent",
                                                     "row_start", "rows"]])
    return h, sizes, flops.value, arrays


def ref_factor_solve(handle, matrix, b=None, kernel=0, threads=1, pivoting=0, max_ratio=1e7, bailout=0,
                     panel_values=0):
    """The reference's factor (+ solve of b in place when given): (status, message, inertia, counts, times, panels,
    diag)."""
    lib = reference()
    inertia = np.zeros(3, np.int64)
    counts = np.zeros(4, np.int64)
    times = np.zeros(2)
    panels = np.zeros(panel_values) if panel_values else None
    diag = np.zeros(matrix.n)
    err = ctypes.create_string_buffer(4096)
    nrhs = 0 if b is None else b.shape[1]
    rc = lib.ref_factor_solve(handle, matrix.n, P(matrix.offsets), P(matrix.rows), P(matrix.values), kernel, threads,
                              pivoting, max_ratio, bailout, P(b), nrhs, P(inertia), P(counts), P(times), P(panels),
                              P(diag), err, 4096)
    return rc, err.value.decode(errors="replace"), inertia, counts, times, panels, diag


def backward_error(matrix, x, b):
    """max over columns of ‖b − A·x‖∞ / (‖A‖∞·‖x‖∞ + ‖b‖∞), the products by the oracle's SpMV (the definition the
    refined solves stop on, computed on the host, independent of the device's own figure)."""
    x = np.asarray(x, dtype=np.float64).reshape(matrix.n, -1)
    b = np.asarray(b, dtype=np.float64).reshape(matrix.n, -1)
    norm = matrix.norm_inf()
    eta = 0.0
    for h in range(x.shape[1]):
        r = b[:, h] - oracle_spmv(matrix, x[:, h])
        denominator = norm * np.max(np.abs(x[:, h])) + np.max(np.abs(b[:, h]))
        eta = max(eta, float(np.max(np.abs(r)) / denominator) if denominator > 0 else float(np.max(np.abs(r))))
    return eta


def ref_combine(alpha, a, beta, b):
    """The reference's SymmetricCsc::combine values of alpha·A + beta·B (A and B of one structure)."""
    out = np.zeros_like(a.values)
    err = ctypes.create_string_buffer(1024)
    rc = reference().ref_combine(a.n, P(a.offsets), P(a.rows), alpha, P(a.values), beta, P(b.values), P(out), err,
                                 1024)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return out


def ref_scale_symmetric(matrix, scale):
    values = np.array(matrix.values, dtype=np.float64, copy=True)
    err = ctypes.create_string_buffer(1024)
    rc = reference().ref_scale_symmetric(matrix.n, P(matrix.offsets), P(matrix.rows), P(values),
                                         P(np.ascontiguousarray(scale, dtype=np.float64)), err, 1024)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return values


def ref_pencil_scale_of(matrix):
    scale = np.zeros(matrix.n)
    err = ctypes.create_string_buffer(1024)
    rc = reference().ref_pencil_scale_of(matrix.n, P(matrix.offsets), P(matrix.rows), P(matrix.values), P(scale),
                                         err, 1024)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return scale


def multifrontal_partial(sym, matrix, max_ratio=1e7):
    """(panels, D per equation, K*, OrcResult) of the oracle's partial factorisation (the last supernode is the
    trailing set: assembled, not eliminated)."""
    lib = oracle()
    t = int(sym.column_count[-1])
    panels = np.zeros(sym.panel_value_count)
    diag = np.zeros(sym.n)
    schur = np.zeros((t, t), order="F")
    res = OrcResult()
    s = orc_symbolic(sym)
    lib.orc_multifrontal_partial(ctypes.byref(s), P(matrix.offsets), P(matrix.rows), P(matrix.values), max_ratio,
                                 P(panels), P(diag), P(schur), ctypes.byref(res))
    return panels, diag, schur, res


def oracle_condense(sym, panels, loads):
    """The trailing rows of the forward sweep over the interior (f*), t × k."""
    x = np.array(loads, dtype=np.float64, order="F", copy=True).reshape(sym.n, -1, order="F")
    s = orc_symbolic(sym)
    oracle().orc_partial_condense(ctypes.byref(s), P(panels), P(x), sym.n, x.shape[1])
    t = int(sym.column_count[-1])
    return np.asfortranarray(x[sym.n - t:, :])


def oracle_recover(sym, panels, loads, trailing_values):
    x = np.array(loads, dtype=np.float64, order="F", copy=True).reshape(sym.n, -1, order="F")
    ub = np.asfortranarray(trailing_values, dtype=np.float64).reshape(-1, x.shape[1], order="F")
    s = orc_symbolic(sym)
    oracle().orc_partial_recover(ctypes.byref(s), P(panels), P(x), sym.n, x.shape[1], P(ub), max(ub.shape[0], 1))
    return x


def ref_symbolic_trailing(graph_offsets, adjacency, weights, order, trailing, amalgamate=1):
    """The reference's SymbolicFactor::analyse of `order` with its last `trailing` positions the trailing set, as
    (handle, sizes, flops, arrays) like ref_symbolic."""
    lib = reference()
    n = int(len(graph_offsets) - 1)
    status = ctypes.c_int()
    err = ctypes.create_string_buffer(1024)
    order = np.ascontiguousarray(order, dtype=np.int32)
    h = lib.ref_analyse_trailing(n, P(graph_offsets), P(adjacency), P(weights), P(order), int(trailing), amalgamate,
                                 ctypes.byref(status), err, 1024)
    if status.value != 0:
        raise RuntimeError(err.value.decode())
    sizes = np.zeros(7, np.int64)
    flops = dbl()
    lib.ref_sym_sizes(h, P(sizes), ctypes.byref(flops))
    S = int(sizes[2])
    arrays = dict(node_order=np.zeros(n, np.int32), first_column=np.zeros(S, np.int32),
                  column_count=np.zeros(S, np.int32), parent=np.zeros(S, np.int32),
                  row_start=np.zeros(S + 1, np.int64), rows=np.zeros(int(sizes[3]), np.int32))
    lib.ref_sym_export(h, *[P(arrays[k]) for k in ["node_order", "first_column", "column_count", "parent",
                                                     "row_start", "rows"]])
    return h, sizes, flops.value, arrays


def ref_partial(handle, matrix, trailing, interior_panel_values, loads=None, trailing_values=None, max_ratio=1e7,
                kernel=0):
    """The reference's PartialFactor: (K*, interior panels, D, inertia, condensed, recovered)."""
    lib = reference()
    n = matrix.n
    schur = np.zeros((trailing, trailing), order="F")
    panels = np.zeros(interior_panel_values)
    diag = np.zeros(n)
    inertia = np.zeros(3, np.int64)
    nrhs = 0 if loads is None else loads.shape[1]
    f = None if loads is None else np.asfortranarray(loads, dtype=np.float64)
    condensed = np.zeros((trailing, nrhs), order="F") if nrhs else None
    ub = None if trailing_values is None else np.asfortranarray(trailing_values, dtype=np.float64)
    recovered = np.zeros((n, nrhs), order="F") if ub is not None else None
    err = ctypes.create_string_buffer(4096)
    rc = lib.ref_partial(handle, n, P(matrix.offsets), P(matrix.rows), P(matrix.values), kernel, max_ratio, P(schur),
                         P(panels), P(diag), P(inertia), P(f), nrhs, P(condensed), P(ub), P(recovered), err, 4096)
    if rc != 0:
        raise RuntimeError(f"ref_partial {rc}: {err.value.decode(errors='replace')}")
    return schur, panels, diag, inertia, condensed, recovered


def oracle_csc_from_blocks(pattern, values, equation_of_dof, kept):
    """(offsets, rows, values) of the oracle's restatement of SymmetricCsc::fromBlocks."""
    lib = oracle()
    values = np.ascontiguousarray(values, dtype=np.float64)
    eq = np.ascontiguousarray(equation_of_dof, dtype=np.int32)
    offsets = np.zeros(kept + 1, np.int64)
    args = (pattern.node_count, P(pattern.first_dof), P(pattern.column_start), P(pattern.block_rows),
            P(pattern.block_value_offsets), P(values), P(eq), kept)
    nnz = lib.orc_csc_from_blocks(*args, P(offsets), None, None)
    rows = np.zeros(max(nnz, 1), np.int32)
    out = np.zeros(max(nnz, 1))
    lib.orc_csc_from_blocks(*args, P(offsets), P(rows), P(out))
    return offsets, rows[:nnz], out[:nnz]


def _check_ref(rc, err):
    if rc != 0:
        raise RuntimeError(f"reference {rc}: {err.value.decode(errors='replace')}")


def ref_pattern(graph_offsets, adjacency, dof_counts):
    """The reference's SparsePattern arrays (first_dof, column_start, block_rows, block_value_offsets)."""
    lib = reference()
    nodes = int(len(graph_offsets) - 1)
    goff = np.ascontiguousarray(graph_offsets, dtype=np.int64)
    adj = np.ascontiguousarray(adjacency, dtype=np.int32)
    counts = np.ascontiguousarray(dof_counts, dtype=np.uint8)
    sizes = np.zeros(2, np.int64)
    err = ctypes.create_string_buffer(1024)
    _check_ref(lib.ref_pattern_export(nodes, P(goff), P(adj), P(counts), P(sizes), None, None, None, None, err, 1024),
               err)
    first = np.zeros(nodes + 1, np.int64)
    cols = np.zeros(nodes + 1, np.int64)
    rows = np.zeros(int(sizes[0]), np.int32)
    voff = np.zeros(int(sizes[0]) + 1, np.int64)
    _check_ref(lib.ref_pattern_export(nodes, P(goff), P(adj), P(counts), P(sizes), P(first), P(cols), P(rows),
                                      P(voff), err, 1024), err)
    return first, cols, rows, voff


def ref_selection(dof_counts, eliminated, node_order):
    lib = reference()
    counts = np.ascontiguousarray(dof_counts, dtype=np.uint8)
    total = int(np.sum(counts.astype(np.int64)))
    elim = None if eliminated is None else np.ascontiguousarray(eliminated, dtype=np.uint8)
    order = np.ascontiguousarray(node_order if node_order is not None else [], dtype=np.int32)
    eq = np.zeros(total, np.int32)
    kept = ctypes.c_int32()
    err = ctypes.create_string_buffer(1024)
    _check_ref(lib.ref_selection(len(counts), P(counts), P(elim), total, P(order), len(order), P(eq),
                                 ctypes.byref(kept), err, 1024), err)
    return eq, kept.value


def ref_from_blocks(graph_offsets, adjacency, dof_counts, values, eliminated, node_order, threads=1):
    lib = reference()
    nodes = int(len(graph_offsets) - 1)
    goff = np.ascontiguousarray(graph_offsets, dtype=np.int64)
    adj = np.ascontiguousarray(adjacency, dtype=np.int32)
    counts = np.ascontiguousarray(dof_counts, dtype=np.uint8)
    total = int(np.sum(counts.astype(np.int64)))
    vals = np.ascontiguousarray(values, dtype=np.float64)
    elim = None if eliminated is None else np.ascontiguousarray(eliminated, dtype=np.uint8)
    order = np.ascontiguousarray(node_order if node_order is not None else [], dtype=np.int32)
    nnz = ctypes.c_int64()
    err = ctypes.create_string_buffer(1024)
    args = (nodes, P(goff), P(adj), P(counts), P(vals), P(elim), total, P(order), len(order), threads,
            ctypes.byref(nnz))
    _check_ref(lib.ref_from_blocks(*args, None, None, None, err, 1024), err)
    kept = int(ref_selection(dof_counts, eliminated, node_order)[1])
    offsets = np.zeros(kept + 1, np.int64)
    rows = np.zeros(max(nnz.value, 1), np.int32)
    out = np.zeros(max(nnz.value, 1))
    _check_ref(lib.ref_from_blocks(*args, P(offsets), P(rows), P(out), err, 1024), err)
    return offsets, rows[:nnz.value], out[:nnz.value]
