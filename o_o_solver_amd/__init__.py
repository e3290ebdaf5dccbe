"""O_o_Solver's numeric multifrontal LDLᵀ (factor and solve) on an AMD Instinct MI355X.

The host-side mirror of the reference's interface for this path (jcdulas/O_o_Solver, include/oos/linalg):

    SymmetricCsc          include/oos/linalg/sparse/SymmetricCsc.hpp:21-108
    SymbolicFactor        include/oos/linalg/symbolic/SymbolicFactor.hpp:37-125 (analysis on the host)
    SupernodalFactor      include/oos/linalg/supernodal/SupernodalFactor.hpp:166-294 (numeric phase on the GPU)
    FactorSolve           include/oos/linalg/supernodal/FactorSolve.hpp:59-87
    FactorOptions, PivotingMode, PivotPolicy, Inertia, PivotDiagnostics

Errors mirror the reference's exception taxonomy: ValueError for std::invalid_argument, MemoryBudgetError,
SingularityError, InternalError; HipError for a failure of the HIP runtime; NotConvergedError when a refined solve
misses its target. All numeric work runs in liboosg.so (hand-written HIP for gfx950); nothing here computes.
"""

import ctypes
import enum
from dataclasses import dataclass, field

import numpy as np

from . import _native

__all__ = ["DeviceMatrix", "SymmetricCsc", "SymbolicFactor", "SupernodalFactor", "FactorSolve", "FactorOptions", "PivotingMode",
           "Precision", "PivotPolicy", "Inertia", "PivotDiagnostics", "PivotIssue", "SolveReport", "Plan",
           "MemoryBudgetError", "MemoryShortfall", "SingularityError", "PrecisionInsufficientError",
           "OperationCancelledError", "InternalError", "HipError", "NotConvergedError", "nested_dissection",
           "estimate_memory_bytes", "PartialFactor", "matrix_checksum", "SparsePattern", "DofSelection",
           "BlockSparseMatrix", "csc_from_blocks_device", "OutOfCoreFactor", "OutOfCoreTraffic"]


@dataclass(frozen=True)
class MemoryShortfall:
    """What a refused reservation asked for (MemoryShortfall, include/oos/base/errors/MemoryBudgetError.hpp:24-29)."""
    requested_bytes: int
    available_bytes: int
    limit_bytes: int


class MemoryBudgetError(MemoryError):
    """The factorisation does not fit the device memory (refused before any numeric work); `shortfall` says by how
    much (MemoryBudgetError.hpp:31-36)."""

    def __init__(self, message, shortfall=None):
        super().__init__(message)
        self.shortfall = shortfall


class SingularityError(RuntimeError):
    """A zero or non-finite pivot in positive definite mode, or a failed ratio test that stops the run; `dofs` lists
    the equations (or the names an equation namer gives them) in elimination order, at most 1000
    (SingularityError::dofs, include/oos/base/errors/SingularityError.hpp:43-50)."""

    def __init__(self, message, dofs=()):
        super().__init__(message)
        self.dofs = list(dofs)


class PrecisionInsufficientError(RuntimeError):
    """An FP32 factor met a zero pivot or failed the ratio test: the matrix is not shown singular (factorise it in
    FP64, which decides). `dofs` as SingularityError's."""

    def __init__(self, message, dofs=()):
        super().__init__(message)
        self.dofs = list(dofs)


class OperationCancelledError(RuntimeError):
    """The cancellation callback asked to stop (OperationCancelledError,
    include/oos/base/errors/OperationCancelledError.hpp)."""


class InternalError(RuntimeError):
    pass


class HipError(RuntimeError):
    pass


class NotConvergedError(RuntimeError):
    pass


_ERRORS = {1: ValueError, 2: MemoryBudgetError, 3: SingularityError, 4: OperationCancelledError, 5: InternalError,
           6: HipError, 7: NotConvergedError, 8: PrecisionInsufficientError}


def _last_equations(lib):
    count = ctypes.c_int64()
    lib.oosg_last_error_equations(None, 0, ctypes.byref(count))
    out = np.zeros(count.value, np.int32)
    lib.oosg_last_error_equations(_ptr(out), count.value, ctypes.byref(count))
    return out.tolist()


def _error(status, namer=None):
    """The exception of a failed call, with the details the library kept for it."""
    lib = _native.lib()
    message = lib.oosg_last_error().decode(errors="replace")
    kind = _ERRORS.get(status, RuntimeError)
    if kind in (SingularityError, PrecisionInsufficientError):
        equations = _last_equations(lib)
        if namer is not None:
            # The reference names equations through FactorOptions::equationNamer (describeEquation,
            # PivotDiagnostics.hpp:122-127): "equation 12" becomes the namer's name of equation 12.
            for e in equations[:20]:
                message = message.replace(f"  equation {e}\n", f"  {namer(e)}\n")
            return kind(message, [namer(e) for e in equations])
        return kind(message, equations)
    if kind is MemoryBudgetError:
        sf = _native.MemoryShortfallC()
        lib.oosg_last_error_shortfall(ctypes.byref(sf))
        return kind(message, MemoryShortfall(int(sf.requested_bytes), int(sf.available_bytes), int(sf.limit_bytes)))
    return kind(message)


def _check(status, namer=None):
    if status != 0:
        raise _error(status, namer)


def _follow_torch(plan, device_tensor):
    """The library's calls wait for the work torch queued on its current stream before reading `device_tensor`."""
    import torch
    stream = torch.cuda.current_stream(device_tensor.device)
    _check(_native.lib().oosg_plan_set_stream(plan.handle, ctypes.c_void_p(stream.cuda_stream)))


def _ptr(array):
    return array.ctypes.data_as(ctypes.c_void_p) if array is not None else None


class PivotingMode(enum.IntEnum):
    POSITIVE_DEFINITE = 0
    INDEFINITE = 1


class Precision(enum.IntEnum):
    FP64 = 0
    FP32 = 1
    # The learned selector chooses the precision of the fronts (FactorOptions.selector, or the committed weights).
    LEARNED = 2
    # FP32 fronts in the lowest FactorOptions.fp32_levels levels of the supernode tree (the subtrees below a cut),
    # FP64 above, on the diagonally scaled matrix; refined by conjugate gradients as FP32 (OOSG_MIXED).
    MIXED = 3


# The C ABI's oosg_precision of each precision a factor is made in.
_C_PRECISION = {Precision.FP64: 0, Precision.FP32: 1, Precision.MIXED: 2}


@dataclass
class PivotPolicy:
    """PARAM MaxDiagonalRatio and SingularityBailOut (include/oos/linalg/supernodal/PivotDiagnostics.hpp:34-44)."""
    max_ratio: float = 1.0e7
    bailout: int = 0


@dataclass
class FactorOptions:
    """FactorOptions (SupernodalFactor.hpp:63-97): the pivoting mode and the pivot policy; `precision` chooses the
    storage of the fronts (FP32 fronts are factorised on the diagonally scaled matrix and refined in FP64)."""
    pivoting: PivotingMode = PivotingMode.POSITIVE_DEFINITE
    pivot_policy: PivotPolicy = field(default_factory=PivotPolicy)
    precision: Precision = Precision.FP64
    # The diagonal the ratio test divides by the pivots: one value per equation in factor order (a negative value
    # keeps the matrix diagonal there), or None (FactorOptions::referenceDiagonal, SupernodalFactor.hpp:75).
    reference_diagonal: object = None
    # Precision.LEARNED: the selector (a selector.PrecisionSelector; None: the committed weights) and the probability
    # from which it takes a cut (the levels below it in FP32) rather than FP64 fronts throughout.
    selector: object = None
    fp32_threshold: float = 0.5
    # Names the equations of the diagnostics (FactorOptions::equationNamer, SupernodalFactor.hpp:76-77): a callable
    # equation -> str; None names them by number.
    equation_namer: object = None
    # Polled between the parts of a call (FactorOptions::cancellation): a callable returning True to stop, or None.
    cancellation: object = None
    # Precision.MIXED: the levels (heights in the supernode tree, from the leaves) factorised in FP32.
    fp32_levels: int = 0


@dataclass(frozen=True)
class Inertia:
    positive: int
    negative: int
    zero: int


@dataclass(frozen=True)
class PivotIssue:
    """One pivot failing the ratio test (PivotIssue, include/oos/linalg/supernodal/PivotDiagnostics.hpp:54-62)."""
    equation: int
    matrix_diagonal: float
    factor_diagonal: float
    ratio: float
    negative_pivot: bool


@dataclass(frozen=True)
class PivotDiagnostics:
    issue_count: int
    first_issue_equation: int
    first_issue_negative: bool
    largest_ratio: float
    largest_ratio_equation: int
    issues: tuple = ()  # the first 1000 PivotIssue, in elimination order


@dataclass(frozen=True)
class SolveReport:
    steps: int
    converged: bool
    backward_error: float
    seconds: float


class SymmetricCsc:
    """The lower triangle by columns, int64 offsets, int32 rows, float64 values (SymmetricCsc.hpp:14-20)."""

    def __init__(self, n, offsets, rows, values):
        self.n = int(n)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.rows = np.ascontiguousarray(rows, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        if self.offsets.shape != (self.n + 1,) or self.rows.shape != self.values.shape or \
                self.offsets[-1] != self.rows.shape[0]:
            raise ValueError("SymmetricCsc: offsets of n + 1 values and one row per value expected")

    @classmethod
    def from_blocks(cls, matrix, selection, device=0):
        """SymmetricCsc::fromBlocks (SymmetricCsc.hpp:40-42), converted on the GPU and copied to the host."""
        offsets, rows, values = csc_from_blocks_device(matrix, selection, device)
        return cls(selection.kept_dof_count, offsets.cpu().numpy(), rows.cpu().numpy(), values.cpu().numpy())

    @classmethod
    def from_compressed_columns(cls, n, offsets, rows, values):
        return cls(n, offsets, rows, values)

    @property
    def non_zero_count(self):
        return int(self.rows.shape[0])

    def desc(self):
        return _native.CscDesc(self.n, _ptr(self.offsets), _ptr(self.rows), _ptr(self.values))

    def multiply(self, x):
        """y = A·x from the lower triangle (SymmetricCsc::multiply, SymmetricCsc.hpp:86-96), host arrays n or n × k."""
        x = np.asarray(x, dtype=np.float64)
        columns = np.repeat(np.arange(self.n), np.diff(self.offsets))
        off = self.rows != columns
        one = x.ndim == 1
        xs = x.reshape(self.n, -1)
        y = np.zeros_like(xs)
        for h in range(xs.shape[1]):
            y[:, h] = np.bincount(self.rows, weights=self.values * xs[columns, h], minlength=self.n)
            y[:, h] += np.bincount(columns[off], weights=self.values[off] * xs[self.rows[off], h], minlength=self.n)
        return y.reshape(-1) if one else y

    def norm_inf(self):
        """‖A‖∞: the largest row sum of |a_ij| over both triangles."""
        columns = np.repeat(np.arange(self.n), np.diff(self.offsets))
        off = self.rows != columns
        sums = np.bincount(self.rows, weights=np.abs(self.values), minlength=self.n)
        sums += np.bincount(columns[off], weights=np.abs(self.values[off]), minlength=self.n)
        return float(np.max(sums)) if self.n else 0.0

    def to_dense(self):
        """The full matrix (tests, small sizes)."""
        dense = np.zeros((self.n, self.n))
        cols = np.repeat(np.arange(self.n), np.diff(self.offsets))
        dense[self.rows, cols] = self.values
        dense[cols, self.rows] = self.values
        return dense


class SymbolicFactor:
    """The symbolic factor of a node graph under an order (SymbolicFactor::analyse): the final (postordered) node
    order, the supernodes, their rows, sizes and flops. Computed on the host by liboosg.so's analysis, which gives
    the reference's structure for the same graph, weights and order."""

    def __init__(self, n, first_column, column_count, parent, row_start, rows, node_order=None, flop_count=None,
                 panel_value_count=None, factor_entry_count=None, trailing_dof_count=0):
        self.n = int(n)
        # The equations of the trailing set (the last supernode when > 0; SymbolicFactor::trailingDofCount).
        self.trailing_dof_count = int(trailing_dof_count)
        self.first_column = np.ascontiguousarray(first_column, dtype=np.int32)
        self.column_count = np.ascontiguousarray(column_count, dtype=np.int32)
        self.parent = np.ascontiguousarray(parent, dtype=np.int32)
        self.row_start = np.ascontiguousarray(row_start, dtype=np.int64)
        self.rows_all = np.ascontiguousarray(rows, dtype=np.int32)
        self.node_order = None if node_order is None else np.ascontiguousarray(node_order, dtype=np.int32)
        c = self.column_count.astype(np.int64)
        r = np.diff(self.row_start)
        self.panel_value_count = int(np.sum(c * (c + r))) if panel_value_count is None else panel_value_count
        self.factor_entry_count = int(np.sum(c * (c + 1) // 2 + c * r)) if factor_entry_count is None \
            else factor_entry_count
        if flop_count is None:
            flop_count = 0.0
            for cc, rr in zip(c.tolist(), r.tolist()):
                m = np.arange(rr, rr + cc, dtype=np.float64)
                flop_count += float(np.sum(m * (m + 1) + 2 * m))
        self.flop_count = flop_count
        self._plans = {}

    @classmethod
    def analyse(cls, graph_offsets, adjacency, node_weights, order, trailing_node_count=0, amalgamate=True):
        lib = _native.lib()
        graph_offsets = np.ascontiguousarray(graph_offsets, dtype=np.int64)
        adjacency = np.ascontiguousarray(adjacency, dtype=np.int32)
        node_weights = np.ascontiguousarray(node_weights, dtype=np.int32)
        order = np.ascontiguousarray(order, dtype=np.int32)
        nodes = int(graph_offsets.shape[0] - 1)
        if node_weights.shape != (nodes,) or order.shape != (nodes,):
            raise ValueError("SymbolicFactor: expected an ordering and one weight per node")
        handle = lib.oosa_analyse(nodes, _ptr(graph_offsets), _ptr(adjacency), _ptr(node_weights), _ptr(order),
                                  int(trailing_node_count), 1 if amalgamate else 0)
        if not handle:
            raise ValueError(lib.oosa_last_error().decode())
        try:
            sizes = np.zeros(8, np.int64)
            flops = ctypes.c_double()
            lib.oosa_sizes(handle, _ptr(sizes), ctypes.byref(flops))
            S = int(sizes[2])
            node_order = np.zeros(nodes, np.int32)
            fc, cc, pa = np.zeros(S, np.int32), np.zeros(S, np.int32), np.zeros(S, np.int32)
            rs, rows = np.zeros(S + 1, np.int64), np.zeros(int(sizes[3]), np.int32)
            lib.oosa_export(handle, _ptr(node_order), _ptr(fc), _ptr(cc), _ptr(pa), _ptr(rs), _ptr(rows))
        finally:
            lib.oosa_free(handle)
        return cls(int(sizes[1]), fc, cc, pa, rs, rows, node_order=node_order, flop_count=flops.value,
                   panel_value_count=int(sizes[4]), factor_entry_count=int(sizes[5]),
                   trailing_dof_count=int(sizes[7]))

    @property
    def supernode_count(self):
        return int(self.first_column.shape[0])

    @property
    def dof_count(self):
        return self.n

    def rows(self, s):
        return self.rows_all[self.row_start[s]:self.row_start[s + 1]]

    def row_count(self, s):
        return int(self.row_start[s + 1] - self.row_start[s])

    def desc(self):
        return _native.SymbolicDesc(self.n, self.supernode_count, _ptr(self.first_column), _ptr(self.column_count),
                                    _ptr(self.parent), _ptr(self.row_start), _ptr(self.rows_all))

    def estimate_memory_bytes(self, pivoting=0):
        """The device bytes a plan of this structure reserves (SupernodalFactor::estimateMemoryBytes)."""
        return estimate_memory_bytes(self, pivoting)

    def plan(self, device=0, hbm_budget_bytes=0, partial=False, out_of_core=False):
        """The device plan of this structure (kept: one per device, budget and kind). partial: the plan of a
        PartialFactor, whose trailing supernode is assembled (K*) and never eliminated; out_of_core: the plan of an
        OutOfCoreFactor, whose panels are spooled to pinned host memory level by level."""
        key = (device, hbm_budget_bytes, bool(partial), bool(out_of_core))
        if key not in self._plans:
            self._plans[key] = Plan(self, device, hbm_budget_bytes, partial, out_of_core)
        return self._plans[key]


class Plan:
    """oosg_plan: levels, relative indices, the kernel schedule and the device memory of one symbolic factor."""

    def __init__(self, symbolic, device=0, hbm_budget_bytes=0, partial=False, out_of_core=False):
        self._lib = _native.lib()
        self.symbolic = symbolic
        handle = ctypes.c_void_p()
        desc = symbolic.desc()
        if partial and out_of_core:
            raise ValueError("plan: a partial plan is made in core")
        if out_of_core:
            _check(self._lib.oosg_plan_create_out_of_core(ctypes.byref(desc), int(device), int(hbm_budget_bytes),
                                                          ctypes.byref(handle)))
        else:
            trailing = symbolic.trailing_dof_count if partial else 0
            _check(self._lib.oosg_plan_create_partial(ctypes.byref(desc), int(trailing), int(device),
                                                      int(hbm_budget_bytes), ctypes.byref(handle)))
        self.partial = bool(partial)
        self.out_of_core = bool(out_of_core)
        self.handle = handle
        self._matrices = {}

    @property
    def levels(self):
        """The plan's levels (heights of the supernode tree)."""
        return int(self.stats()["levels"])

    def stats(self):
        s = np.zeros(8, np.int64)
        _check(self._lib.oosg_plan_stats(self.handle, _ptr(s)))
        keys = ["supernodes", "levels", "panel_values", "update_values", "small_fronts", "large_fronts", "launches",
                "tiles"]
        return dict(zip(keys, s.tolist()))

    PROFILE_KINDS = ["small_fronts", "assembly", "diagonal_blocks", "panel_solve", "panel_update_tiles",
                     "update_matrix_tiles", "forward_triangles", "backward_triangles", "residual", "forward_rows",
                     "backward_columns", "triangle_products", "triangle_inverses", "update_matrix_big_tiles",
                     "panel_update_fused_tiles"]

    def profile(self, enable=True):
        """Per-kernel device times by HIP events on the plan's stream (enabling resets the sums)."""
        _check(self._lib.oosg_plan_profile(self.handle, 1 if enable else 0))

    def profile_read(self):
        """{kind: (seconds, launches, work)}: work in flops for the factor kernels, in bytes for the sweeps and the
        residuals."""
        kinds = len(self.PROFILE_KINDS)
        seconds = np.zeros(kinds)
        launches = np.zeros(kinds, np.int64)
        work = np.zeros(kinds)
        _check(self._lib.oosg_plan_profile_read(self.handle, _ptr(seconds), _ptr(launches), _ptr(work)))
        return {k: (float(seconds[i]), int(launches[i]), float(work[i])) for i, k in enumerate(self.PROFILE_KINDS)}

    def matrix(self, matrix):
        """The device copy of `matrix` (kept while `matrix` is alive and unchanged)."""
        return DeviceMatrix(self, matrix)

    def synchronize(self):
        _check(self._lib.oosg_plan_synchronize(self.handle))

    def memory(self):
        """(bytes of the budget in use: the plan's own memory and its live factors', the budget)."""
        reserved, budget = ctypes.c_uint64(), ctypes.c_uint64()
        _check(self._lib.oosg_plan_memory(self.handle, ctypes.byref(reserved), ctypes.byref(budget)))
        return int(reserved.value), int(budget.value)

    def set_cancellation(self, cancelled):
        """Polls `cancelled()` between the parts of the plan's calls (None: no polling); True stops the call with
        OperationCancelledError once the work queued is done."""
        if cancelled is None:
            self._cancel_fn = None
            _check(self._lib.oosg_plan_set_cancellation(self.handle, _native.CANCEL_FN(), None))
            return
        self._cancel_fn = _native.CANCEL_FN(lambda _context: 1 if cancelled() else 0)
        _check(self._lib.oosg_plan_set_cancellation(self.handle, self._cancel_fn, None))

    def use_graphs(self, enable=True):
        """Replays of the factorisation's and the sweeps' launch sequences as HIP graphs (default on)."""
        _check(self._lib.oosg_plan_use_graphs(self.handle, 1 if enable else 0))

    def tile_policy(self, policy):
        """The update-matrix tile size: 0 64 × 64, 1 128 × 128, 2 by level (the autotuner's choice; the default)."""
        _check(self._lib.oosg_plan_tile_policy(self.handle, int(policy)))

    def launch_times(self):
        """The factorisation's launches: (profile kind, level, ms in the last autotuner run, flops) arrays."""
        count = ctypes.c_int32(0)
        _check(self._lib.oosg_plan_launch_times(self.handle, ctypes.byref(count), None, None, None, None))
        n = count.value
        kinds, levels = np.zeros(n, np.int32), np.zeros(n, np.int32)
        ms, flops = np.zeros(n, np.float32), np.zeros(n)
        _check(self._lib.oosg_plan_launch_times(self.handle, ctypes.byref(count), _ptr(kinds), _ptr(levels), _ptr(ms),
                                                _ptr(flops)))
        return kinds, levels, ms, flops

    def streams(self, precision, count):
        """The factorisation's streams in `precision`: 1 (default) or 3 (trailing updates and whole inverses beside
        the critical path)."""
        _check(self._lib.oosg_plan_streams(self.handle, int(Precision(precision)), int(count)))

    LAUNCH_MODES = ["graph", "graph_streams", "eager_streams"]

    def launch_mode(self, precision, mode=None):
        """How the factorisation in `precision` is launched: "graph" (one HIP graph on the plan's stream, the
        default), "graph_streams" (one graph of three streams) or "eager_streams" (three streams, one launch at a
        time). Sets the mode when given; returns the mode in force."""
        precision = int(Precision(precision))
        if mode is not None:
            _check(self._lib.oosg_plan_launch_mode(self.handle, precision, self.LAUNCH_MODES.index(mode)))
        out = ctypes.c_int32()
        _check(self._lib.oosg_plan_get_launch_mode(self.handle, precision, ctypes.byref(out)))
        return self.LAUNCH_MODES[out.value]

    def autotune(self, device_matrix, precision=None):
        """Times both tile sizes of every update-matrix launch on `device_matrix` (a matrix of this plan) and keeps
        the faster per level, then the whole factorisation in each launch mode, keeping the fastest. Returns
        (launches that chose 128 × 128 tiles, launches measured, streams kept); launch_mode(precision) gives the mode
        kept."""
        precision = Precision.FP64 if precision is None else Precision(precision)
        big, tuned, streams = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib.oosg_plan_autotune(self.handle, device_matrix.handle, int(precision), ctypes.byref(big),
                                            ctypes.byref(tuned), ctypes.byref(streams)))
        return big.value, tuned.value, streams.value

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.oosg_plan_destroy(self.handle)
            self.handle = None


class DeviceMatrix:
    """oosg_matrix: a SymmetricCsc in device memory, with its front positions and its row form for residuals."""

    def __init__(self, plan, matrix):
        self._lib = _native.lib()
        self.plan = plan
        self.matrix = matrix
        self.structure = matrix  # the offsets and rows (the values of `matrix` may be superseded)
        handle = ctypes.c_void_p()
        desc = matrix.desc()
        _check(self._lib.oosg_matrix_create(plan.handle, ctypes.byref(desc), ctypes.byref(handle)))
        self.handle = handle

    @classmethod
    def from_device_csc(cls, plan, offsets, rows, values):
        """oosg_matrix_create_device: a matrix of the plan from a CSC in device memory (device tensors, as
        csc_from_blocks_device makes them); its structure is read once to the host for the plan's positions."""
        self = cls.__new__(cls)
        self._lib = _native.lib()
        self.plan = plan
        n = int(offsets.numel()) - 1
        handle = ctypes.c_void_p()
        _follow_torch(plan, values)
        _check(self._lib.oosg_matrix_create_device(plan.handle, n, offsets.data_ptr(), rows.data_ptr(),
                                                   values.data_ptr(), ctypes.byref(handle)))
        self.handle = handle
        self.structure = SymmetricCsc(n, offsets.cpu().numpy(), rows.cpu().numpy(), values.cpu().numpy())
        self.matrix = None
        return self

    def set_values(self, values):
        values = np.ascontiguousarray(values, dtype=np.float64)
        if values.shape != self.structure.values.shape:
            raise ValueError("DeviceMatrix.set_values: one value per stored entry expected")
        _check(self._lib.oosg_matrix_set_values(self.handle, _ptr(values)))
        self.matrix = SymmetricCsc(self.structure.n, self.structure.offsets, self.structure.rows, values)

    def set_combination(self, alpha, a, beta, b, scale=None):
        """Values S·(alpha·A + beta·B)·S from two device matrices of this plan and structure, on the device
        (SymmetricCsc::combine then scaleSymmetric, the reference's rounding). The host copy `matrix` is then None:
        the values live on the device."""
        scale_arr = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
        if scale_arr is not None and scale_arr.shape != (self.plan.symbolic.n,):
            raise ValueError("SymmetricCsc::scaleSymmetric: one scale per equation expected")
        _check(self._lib.oosg_matrix_set_combination(self.handle, float(alpha), a.handle, float(beta), b.handle,
                                                     _ptr(scale_arr)))
        self.matrix = None

    def set_values_device(self, values):
        """New values from a device tensor in the order of this structure's CSC (a matrix of the same pattern and
        selection converted on the GPU: csc_from_blocks_device), with no host copy (oosg_matrix_set_values_device).
        The host copy `matrix` is then None."""
        if values.numel() != self.structure.values.shape[0]:
            raise ValueError("DeviceMatrix.set_values_device: one value per stored entry expected")
        import torch
        values = values.to(torch.float64).contiguous()
        _follow_torch(self.plan, values)
        _check(self._lib.oosg_matrix_set_values_device(self.handle, values.data_ptr()))
        self.matrix = None

    def multiply(self, x):
        """y = A·x for a torch tensor on the plan's device, n × k (column-major storage: x.T contiguous)."""
        import torch
        xs = _as_device_columns(x)
        y = torch.empty_like(xs)
        _follow_torch(self.plan, xs)
        _check(self._lib.oosg_matrix_multiply_device(self.handle, xs.data_ptr(), y.data_ptr(), xs.shape[1],
                                                     xs.shape[0]))
        return y.T if x.dim() == 2 else y.reshape(-1)

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.oosg_matrix_destroy(self.handle)
            self.handle = None


def _as_device_columns(x):
    """A (k, n) contiguous float64 view of an n × k (or n) tensor: column j of x is row j."""
    if x.dim() == 1:
        return x.reshape(1, -1).contiguous().double()
    return x.T.contiguous().double()


class SupernodalFactor:
    """The numeric factorisation of a SymmetricCsc in the order of its symbolic factor, on the GPU."""

    def __init__(self, handle, plan, device_matrix, options):
        self._lib = _native.lib()
        self.handle = handle
        self.plan = plan
        self.device_matrix = device_matrix
        self.options = options
        self.selected_probability = None
        # Precision.LEARNED: whether the FP32 factor it chose failed and FP64 took over.
        self.fell_back = False
        info = _native.FactorInfoC()
        _check(self._lib.oosg_factor_info_get(handle, ctypes.byref(info)))
        self._info = info

    @classmethod
    def _call(cls, symbolic, matrix, options, device, hbm_budget_bytes, device_matrix, plan=None):
        """The plan, the device matrix, the C options (with the precision the selector picks, for LEARNED), the
        reference diagonal they point to, and the selector's probability."""
        options = options or FactorOptions()
        plan = plan or (device_matrix.plan if device_matrix is not None else symbolic.plan(device, hbm_budget_bytes))
        if device_matrix is None:
            device_matrix = plan.matrix(matrix)
        reference = None
        if options.reference_diagonal is not None:
            reference = np.ascontiguousarray(options.reference_diagonal, dtype=np.float64)
            if reference.shape != (symbolic.n,):
                # requireReferenceDiagonal (src/linalg/supernodal/PivotDiagnostics.cpp:24-32)
                raise ValueError(f"factorisation: the reference diagonal of the ratio test has {reference.size} "
                                 f"values and the matrix {symbolic.n} equations; give one value per equation "
                                 "(a negative value where the matrix diagonal is kept) or none")
        precision = options.precision
        selected = None
        fp32_levels = options.fp32_levels
        if precision == Precision.LEARNED:
            # The selector's cut: the levels below it in FP32 (all of them: FP32 fronts; none: FP64 fronts).
            from .selector import PrecisionSelector
            selector = options.selector or PrecisionSelector.trained(device)
            if selector is None:
                raise ValueError("factorise: Precision.LEARNED needs a selector (no committed weights found)")
            cut, probabilities = selector.choose(plan, device_matrix, options.fp32_threshold)
            selected = float(probabilities.max()) if probabilities.size else 0.0
            if cut == 0:
                precision = Precision.FP64
            elif cut >= probabilities.size:
                precision = Precision.FP32
            else:
                precision, fp32_levels = Precision.MIXED, cut
        opts = cls._options_c(options, precision, reference, fp32_levels)
        plan.set_cancellation(options.cancellation)
        return options, plan, device_matrix, opts, reference, selected

    @staticmethod
    def _options_c(options, precision, reference, fp32_levels=None):
        levels = options.fp32_levels if fp32_levels is None else fp32_levels
        return _native.FactorOptionsC(int(options.pivoting), _C_PRECISION[Precision(precision)],
                                      float(options.pivot_policy.max_ratio), int(options.pivot_policy.bailout),
                                      _ptr(reference), int(levels))

    @classmethod
    def factorise(cls, symbolic, matrix, options=None, device=0, hbm_budget_bytes=0, device_matrix=None, plan=None):
        """SupernodalFactor::factorise (SupernodalFactor.hpp:176-183). `matrix` is a SymmetricCsc in the factor's
        order; `device_matrix` (optional) is its device copy, reused across factorisations. With Precision.LEARNED,
        an FP32 factor that fails its pivot tests is replaced by the FP64 factor (`fell_back`)."""
        options, plan, device_matrix, opts, reference, selected = cls._call(symbolic, matrix, options, device,
                                                                            hbm_budget_bytes, device_matrix, plan)
        handle = ctypes.c_void_p()
        lib = _native.lib()
        status = lib.oosg_factorise(plan.handle, device_matrix.handle, ctypes.byref(opts), ctypes.byref(handle))
        fell_back = False
        if status == 8 and options.precision == Precision.LEARNED:
            opts = cls._options_c(options, Precision.FP64, reference)
            status = lib.oosg_factorise(plan.handle, device_matrix.handle, ctypes.byref(opts), ctypes.byref(handle))
            fell_back = True
        _check(status, options.equation_namer)
        factor = cls(handle, plan, device_matrix, options)
        factor.selected_probability = selected
        factor.fell_back = fell_back
        return factor

    @classmethod
    def factorise_and_solve(cls, symbolic, matrix, b, options=None, device=0, hbm_budget_bytes=0,
                            device_matrix=None, target=1e-10, max_steps=10):
        """factorise, then FactorSolve.solve_refined of the device tensor b (n × k), as one call
        (oosg_factorise_solve_device): the first solve's forward sweep of the lower levels runs beside the
        factorisation of the top levels. Returns (factor, x, SolveReport); NotConvergedError carries the factor as
        its `factor` attribute. With Precision.LEARNED, an FP32 factor that fails its pivot tests or whose refinement
        misses the target is replaced by the FP64 call (`fell_back` on the factor)."""
        import torch
        options, plan, device_matrix, opts, reference, selected = cls._call(symbolic, matrix, options, device,
                                                                            hbm_budget_bytes, device_matrix)
        bc = _as_device_columns(b)
        xc = torch.empty_like(bc)
        lib = _native.lib()

        def run(c_options):
            handle = ctypes.c_void_p()
            report = _native.SolveReportC()
            _follow_torch(plan, bc)
            status = lib.oosg_factorise_solve_device(plan.handle, device_matrix.handle, ctypes.byref(c_options),
                                                     bc.data_ptr(), xc.data_ptr(), bc.shape[1], bc.shape[0],
                                                     float(target), int(max_steps), ctypes.byref(handle),
                                                     ctypes.byref(report))
            factor = cls(handle, plan, device_matrix, options) if handle.value else None
            rep = SolveReport(int(report.steps), bool(report.converged), float(report.backward_error),
                              float(report.seconds))
            return status, factor, rep

        status, factor, rep = run(opts)
        fell_back = False
        scaled = opts.precision != _C_PRECISION[Precision.FP64]
        if status in (7, 8) and options.precision == Precision.LEARNED and scaled:
            factor = None
            status, factor, rep = run(cls._options_c(options, Precision.FP64, reference))
            fell_back = True
        if factor is not None:
            factor.selected_probability = selected
            factor.fell_back = fell_back
        if status == 7:
            error = NotConvergedError(f"{lib.oosg_last_error().decode()} ({rep})")
            error.factor = factor
            raise error
        _check(status, options.equation_namer)
        return factor, (xc.T if b.dim() == 2 else xc.reshape(-1)), rep

    @property
    def dimension(self):
        return self.plan.symbolic.n

    @property
    def inertia(self):
        return Inertia(int(self._info.positive), int(self._info.negative), int(self._info.zero))

    @property
    def is_singular(self):
        return self._info.zero > 0

    @property
    def delayed_pivot_count(self):
        return int(self._info.delayed_pivots)

    @property
    def two_by_two_pivot_count(self):
        return int(self._info.two_by_two_pivots)

    @property
    def pivot_diagnostics(self):
        i = self._info
        recorded = ctypes.c_int64()
        _check(self._lib.oosg_factor_issues(self.handle, None, 0, ctypes.byref(recorded)))
        raw = (_native.PivotIssueC * max(recorded.value, 1))()
        _check(self._lib.oosg_factor_issues(self.handle, raw, recorded.value, ctypes.byref(recorded)))
        issues = tuple(PivotIssue(int(r.equation), float(r.matrix_diagonal), float(r.factor_diagonal), float(r.ratio),
                                  bool(r.kind == 1)) for r in raw[:recorded.value])
        return PivotDiagnostics(int(i.issue_count), int(i.first_issue_equation), bool(i.first_issue_negative),
                                float(i.largest_ratio), int(i.largest_ratio_equation), issues)

    @property
    def factor_seconds(self):
        return float(self._info.factor_seconds)

    @property
    def flop_count(self):
        return float(self._info.flops)

    @property
    def precision(self):
        return {0: Precision.FP64, 1: Precision.FP32, 2: Precision.MIXED}[int(self._info.precision)]

    @property
    def fp32_levels(self):
        """Precision.MIXED: the levels factorised in FP32."""
        return int(self._info.fp32_levels)

    def panel(self, s):
        """(values m × e, rows) of supernode s, in the reference's FactorPanel layout: m = c + r rows, or c + d + r when
        its children delayed d columns into its front (indefinite mode); e eliminated columns."""
        m, e = ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib.oosg_factor_panel_shape(self.handle, int(s), ctypes.byref(m), ctypes.byref(e)))
        m, e = m.value, e.value
        values = np.zeros(m * e)
        rows = np.zeros(m, np.int32)
        eliminated = ctypes.c_int32()
        _check(self._lib.oosg_factor_panel(self.handle, int(s), _ptr(values), _ptr(rows), ctypes.byref(eliminated)))
        return values.reshape(e, m).T, rows

    def subdiagonal(self, s):
        """The subdiagonal of D over the e pivots of supernode s (FactorPanel::subdiagonal): e_k of a 2 × 2 block
        starting at k, 0 elsewhere."""
        m, e = ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib.oosg_factor_panel_shape(self.handle, int(s), ctypes.byref(m), ctypes.byref(e)))
        out = np.zeros(e.value)
        _check(self._lib.oosg_factor_subdiagonal(self.handle, int(s), _ptr(out)))
        return out

    def diagonal(self):
        d = np.zeros(self.dimension)
        _check(self._lib.oosg_factor_diagonal(self.handle, _ptr(d)))
        return d

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.oosg_factor_destroy(self.handle)
            self.handle = None


class FactorSolve:
    """FactorSolve (FactorSolve.hpp:59-87): solves with a factor, in place."""

    def __init__(self, factor):
        self.factor = factor
        self._lib = _native.lib()

    def solve(self, b):
        """X = A⁻¹·B. `b`: a numpy array (n or n × k; solved in place when it is Fortran-ordered float64, a solved
        copy is returned otherwise) or a torch tensor on the GPU (a solved copy is returned)."""
        n = self.factor.dimension
        if hasattr(b, "data_ptr"):
            cols = _as_device_columns(b)
            if cols.shape[1] != n:
                raise ValueError(f"FactorSolve::solve: the right-hand sides have {cols.shape[1]} rows, the factor {n}")
            _follow_torch(self.factor.plan, cols)
            _check(self._lib.oosg_solve_device(self.factor.handle, cols.data_ptr(), n, cols.shape[0]))
            return cols.T if b.dim() == 2 else cols.reshape(-1)
        arr = np.asarray(b)
        vector = arr.ndim == 1
        x = np.array(arr.reshape(n, -1) if not vector else arr.reshape(n, 1), dtype=np.float64, order="F", copy=True)
        if x.shape[0] != n:
            raise ValueError(f"FactorSolve::solve: the right-hand sides have {x.shape[0]} rows, the factor {n}")
        _check(self._lib.oosg_solve(self.factor.handle, _ptr(x), n, x.shape[1]))
        return x.reshape(-1) if vector else x

    def solve_refined(self, b, target=1e-10, max_steps=10, x=None):
        """x = A⁻¹·b by refinement in FP64 with the factorised matrix, to the backward error `target`. `b`: torch
        tensor n × k on the device. Returns (x, SolveReport); raises NotConvergedError when the target is missed."""
        import torch
        bc = _as_device_columns(b)
        xc = torch.empty_like(bc) if x is None else x
        report = _native.SolveReportC()
        _follow_torch(self.factor.plan, bc)
        status = self._lib.oosg_solve_refined_device(self.factor.handle, self.factor.device_matrix.handle,
                                                     bc.data_ptr(), xc.data_ptr(), bc.shape[1], bc.shape[0],
                                                     float(target), int(max_steps), ctypes.byref(report))
        rep = SolveReport(int(report.steps), bool(report.converged), float(report.backward_error),
                          float(report.seconds))
        if status == 7:
            raise NotConvergedError(f"{_native.lib().oosg_last_error().decode()} ({rep})")
        _check(status)
        return (xc.T if b.dim() == 2 else xc.reshape(-1)), rep


class SparsePattern:
    """The node-block pattern of a symmetric matrix (SparsePattern, include/oos/linalg/sparse/SparsePattern.hpp:13-68):
    node n owns the DOFs [first_dof[n], first_dof[n + 1]); block column j is [column_start[j], column_start[j + 1]),
    its diagonal block first, then the neighbours i > j ascending (block_rows); block b is dofs(i) × dofs(j),
    column-major at block_value_offsets[b]. Built as the reference builds it from a node graph
    (src/linalg/sparse/SparsePattern.cpp:10-60)."""

    def __init__(self, first_dof, column_start, block_rows, block_value_offsets):
        self.first_dof = np.ascontiguousarray(first_dof, dtype=np.int64)
        self.column_start = np.ascontiguousarray(column_start, dtype=np.int64)
        self.block_rows = np.ascontiguousarray(block_rows, dtype=np.int32)
        self.block_value_offsets = np.ascontiguousarray(block_value_offsets, dtype=np.int64)

    @classmethod
    def from_graph(cls, graph_offsets, adjacency, node_dof_counts):
        counts = np.asarray(node_dof_counts, dtype=np.int64)
        nodes = counts.shape[0]
        if np.any((counts < 1) | (counts > 6)):
            raise ValueError("SparsePattern: expected 1 to 6 DOFs per node")
        first_dof = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        column_start = [0]
        rows, offsets = [], [0]
        for j in range(nodes):
            neighbours = np.sort(np.asarray(adjacency[graph_offsets[j]:graph_offsets[j + 1]]))
            below = neighbours[neighbours > j]
            for i in [j] + below.tolist():
                rows.append(i)
                offsets.append(offsets[-1] + int(counts[j]) * int(counts[i]))
            column_start.append(len(rows))
        return cls(first_dof, column_start, rows, offsets)

    @property
    def node_count(self):
        return int(self.first_dof.shape[0] - 1)

    @property
    def dof_count(self):
        return int(self.first_dof[-1])

    @property
    def value_count(self):
        return int(self.block_value_offsets[-1])


class DofSelection:
    """The equations of the kept DOFs (DofSelection, include/oos/linalg/sparse/DofSelection.hpp:13-60): numbered node by
    node in `node_order` (empty: natural), by increasing component within a node, eliminated DOFs removed
    (src/linalg/sparse/DofSelection.cpp:22-95). equation_of_dof: per node-major DOF, -1 when eliminated."""

    ELIMINATED = -1

    def __init__(self, node_dof_counts, is_eliminated=None, node_order=None):
        counts = np.asarray(node_dof_counts, dtype=np.int64)
        if np.any((counts < 1) | (counts > 6)):
            raise ValueError("DofSelection: expected 1 to 6 DOFs per node; give the DOF counts of the pattern "
                             "(SparsePattern)")
        first = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        total = int(first[-1])
        kept_flag = np.ones(total, bool) if is_eliminated is None or len(is_eliminated) == 0 else \
            (np.asarray(is_eliminated) == 0)
        if kept_flag.shape[0] != total:
            raise ValueError(f"DofSelection: expected one elimination flag per DOF ({total}), found "
                             f"{kept_flag.shape[0]}")
        nodes = counts.shape[0]
        order = np.arange(nodes) if node_order is None or len(node_order) == 0 else np.asarray(node_order)
        if len(order) > nodes or len(np.unique(order)) != len(order) or np.any((order < 0) | (order >= nodes)):
            raise ValueError(f"DofSelection: expected distinct nodes in [0, {nodes}) in the node order")
        equation = np.full(total, -1, np.int32)
        nxt = 0
        for node in order.tolist():
            dofs = np.arange(first[node], first[node + 1])
            kept = dofs[kept_flag[dofs]]
            equation[kept] = np.arange(nxt, nxt + kept.shape[0], dtype=np.int32)
            nxt += kept.shape[0]
        if nxt != int(np.count_nonzero(kept_flag)):
            raise ValueError("DofSelection: the node order leaves out nodes that keep DOFs")
        self.first_dof = first
        self.equation_of_dof = equation
        self.kept_dof_count = nxt


class BlockSparseMatrix:
    """Values on a SparsePattern (BlockSparseMatrix, include/oos/linalg/sparse/BlockSparseMatrix.hpp:14-36): a numpy
    array or a device tensor of pattern.value_count values."""

    def __init__(self, pattern, values):
        self.pattern = pattern
        self.values = values


def _device_blocks(matrix, selection, device):
    import torch
    dev = torch.device("cuda", device)
    pat = matrix.pattern
    arrays = dict(first_dof=torch.as_tensor(pat.first_dof, device=dev),
                  column_start=torch.as_tensor(pat.column_start, device=dev),
                  block_rows=torch.as_tensor(pat.block_rows, device=dev),
                  block_value_offsets=torch.as_tensor(pat.block_value_offsets, device=dev),
                  values=torch.as_tensor(matrix.values, dtype=torch.float64, device=dev).contiguous(),
                  equation_of_dof=torch.as_tensor(selection.equation_of_dof, device=dev))
    desc = _native.BlockMatrixDesc(pat.node_count, arrays["first_dof"].data_ptr(), arrays["column_start"].data_ptr(),
                                   arrays["block_rows"].data_ptr(), arrays["block_value_offsets"].data_ptr(),
                                   arrays["values"].data_ptr())
    return arrays, desc


def csc_from_blocks_device(matrix, selection, device=0):
    """SymmetricCsc::fromBlocks (SymmetricCsc.hpp:40-42) on the GPU (oosg_csc_from_blocks_device): the lower CSC of the
    kept DOFs in equation order as device tensors (offsets int64, rows int32, values float64), bit-identical to the
    reference's conversion."""
    import torch
    arrays, desc = _device_blocks(matrix, selection, device)
    dev = torch.device("cuda", device)
    lib = _native.lib()
    kept = selection.kept_dof_count
    offsets = torch.empty(kept + 1, dtype=torch.int64, device=dev)
    nnz = ctypes.c_int64()
    stream = torch.cuda.current_stream(dev).cuda_stream
    dofs = int(selection.equation_of_dof.shape[0])
    _check(lib.oosg_csc_from_blocks_device(device, ctypes.byref(desc), arrays["equation_of_dof"].data_ptr(), dofs,
                                           kept, offsets.data_ptr(), None, None, ctypes.byref(nnz),
                                           ctypes.c_void_p(stream)))
    rows = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)
    values = torch.empty(max(nnz.value, 1), dtype=torch.float64, device=dev)
    _check(lib.oosg_csc_from_blocks_device(device, ctypes.byref(desc), arrays["equation_of_dof"].data_ptr(), dofs,
                                           kept, offsets.data_ptr(), rows.data_ptr(), values.data_ptr(),
                                           ctypes.byref(nnz), ctypes.c_void_p(stream)))
    return offsets, rows[:nnz.value], values[:nnz.value]


@dataclass(frozen=True)
class OutOfCoreTraffic:
    """What the spool holds and what the factorisation wrote and the solves read back (OutOfCoreTraffic,
    include/oos/linalg/supernodal/OutOfCoreFactor.hpp:14-20)."""
    spooled_bytes: int
    written_bytes: int
    read_bytes: int
    read_requests: int


class OutOfCoreFactor:
    """OutOfCoreFactor (include/oos/linalg/supernodal/OutOfCoreFactor.hpp:22-95) on the GPU: the factorisation on an
    out-of-core plan (oosg_plan_create_out_of_core) keeps each level's panels and inverses in one of two device windows
    and spools them to pinned host memory once the level is done; the solves copy the levels back, the next level's
    copy beside this level's sweep. The factor and its solves are the in-core factor's bits. Positive definite mode
    (FP64 or FP32 fronts): the reference's indefinite mode of this class is not built here."""

    @classmethod
    def factorise(cls, symbolic, matrix, options=None, device=0, hbm_budget_bytes=0):
        """The factorisation with its panels spooled. ValueError for indefinite pivoting; otherwise what
        SupernodalFactor.factorise raises."""
        options = options or FactorOptions()
        if options.pivoting != PivotingMode.POSITIVE_DEFINITE:
            raise ValueError("OutOfCoreFactor: the out-of-core factorisation runs in positive definite mode here")
        self = cls.__new__(cls)
        plan = symbolic.plan(device, hbm_budget_bytes, out_of_core=True)
        self._factor = SupernodalFactor.factorise(symbolic, matrix, options, device, hbm_budget_bytes,
                                                  device_matrix=plan.matrix(matrix), plan=plan)
        return self

    @property
    def factor(self):
        """The factor (its panels are read from the spool: panel(s), diagonal(), inertia, diagnostics)."""
        return self._factor

    @property
    def inertia(self):
        return self._factor.inertia

    @property
    def is_singular(self):
        return self._factor.is_singular

    def traffic(self):
        t = _native.OutOfCoreTrafficC()
        _check(_native.lib().oosg_factor_out_of_core_traffic(self._factor.handle, ctypes.byref(t)))
        return OutOfCoreTraffic(int(t.spooled_bytes), int(t.written_bytes), int(t.read_bytes), int(t.read_requests))

    @property
    def spooled_bytes(self):
        return self.traffic().spooled_bytes

    def solve(self, b):
        """B ← A⁻¹·B (FactorSolve::solve's bits), as FactorSolve.solve."""
        return FactorSolve(self._factor).solve(b)


def matrix_checksum(matrix):
    """The 64-bit FNV-1a checksum of a SymmetricCsc's structure and values (oosg_csc_checksum), by which a recomputed
    interior factor is refused for another matrix (PartialFactor.cpp:40-60)."""
    out = ctypes.c_uint64()
    desc = matrix.desc()
    _check(_native.lib().oosg_csc_checksum(ctypes.byref(desc), ctypes.byref(out)))
    return int(out.value)


class PartialFactor:
    """PartialFactor (include/oos/linalg/supernodal/PartialFactor.hpp:47-150) on the GPU: the interior equations of a
    unit eliminated, the trailing set (its boundary, ordered last by the symbolic factor: SymbolicFactor.analyse with
    trailing_node_count) left as the exact Schur complement

        K* = K_bb − K_bi·K_ii⁻¹·K_ib,

    with the same elimination applied to loads (condense_loads: f* = f_b − K_bi·K_ii⁻¹·f_i) and the back-substitution
    of the interior from given boundary values (recover: u_i = K_ii⁻¹·(f_i − K_ib·u_b)). The interior is factorised in
    positive definite mode and FP64 on the partial plan of the symbolic factor (oosg_plan_create_partial), whose
    trailing front is assembled into K* after the last level. The interior factor stays in device memory while kept;
    release_interior_factor frees it and recompute_interior_factor computes it again, to the same bits (the
    factorisation is deterministic)."""

    @classmethod
    def factorise(cls, symbolic, matrix, options=None, device=0, hbm_budget_bytes=0):
        """PartialFactor::factorise (PartialFactor.cpp:65-83). ValueError when options.pivoting is not positive
        definite; MemoryBudgetError before any work; SingularityError for a mechanism of the interior."""
        options = options or FactorOptions()
        cls._require_positive_definite(options)
        self = cls.__new__(cls)
        self._dimension = symbolic.dof_count
        self._trailing = symbolic.trailing_dof_count
        self._device = device
        self._budget = hbm_budget_bytes
        self._interior_supernodes = symbolic.supernode_count - (1 if self._trailing > 0 else 0)
        self._interior = self._factorise_interior(symbolic, matrix, options)
        self._pivot_policy = PivotPolicy(options.pivot_policy.max_ratio, options.pivot_policy.bailout)
        self._checksum = matrix_checksum(matrix)
        t = self._trailing
        self._schur = np.zeros((t, t), order="F")
        if t > 0:
            _check(_native.lib().oosg_partial_schur_complement(self._interior.handle, _ptr(self._schur), t))
        return self

    @staticmethod
    def _require_positive_definite(options):
        if options.pivoting != PivotingMode.POSITIVE_DEFINITE:
            raise ValueError("PartialFactor: the interior of a unit is factorised in positive definite mode (a "
                             "delayed pivot would move an interior equation into the boundary); pass "
                             "PivotingMode::PositiveDefinite")

    def _factorise_interior(self, symbolic, matrix, options):
        plan = symbolic.plan(self._device, self._budget, partial=self._trailing > 0)
        device_matrix = plan.matrix(matrix)
        reference = None
        if options.reference_diagonal is not None:
            reference = np.ascontiguousarray(options.reference_diagonal, dtype=np.float64)
        opts = SupernodalFactor._options_c(options, Precision.FP64, reference)
        plan.set_cancellation(options.cancellation)
        handle = ctypes.c_void_p()
        _check(_native.lib().oosg_factorise(plan.handle, device_matrix.handle, ctypes.byref(opts),
                                            ctypes.byref(handle)), options.equation_namer)
        return SupernodalFactor(handle, plan, device_matrix, options)

    @property
    def dimension(self):
        return self._dimension

    @property
    def trailing_dof_count(self):
        return self._trailing

    @property
    def interior_dof_count(self):
        return self._dimension - self._trailing

    def schur_complement(self):
        """K*, trailing_dof_count square, both triangles, in the order of the trailing equations."""
        return self._schur

    @property
    def is_interior_factor_kept(self):
        return self._interior is not None

    @property
    def pivot_policy(self):
        return self._pivot_policy

    def _require_kept(self, operation):
        if self._interior is None:
            raise RuntimeError(f"PartialFactor::{operation}: the interior factor was released; call "
                               "recomputeInteriorFactor() first (the same bits as the first factorisation)")

    def _columns(self, operation, array, expected, what):
        a = np.asarray(array)
        rows = a.shape[0] if a.ndim >= 1 else 0
        if rows != expected:
            raise ValueError(f"PartialFactor::{operation}: {what} have {rows} rows, expected {expected}")
        return a.reshape(expected, -1) if a.ndim == 1 else a

    @property
    def interior_factor(self):
        """The interior factor (its inertia, its ratio test)."""
        self._require_kept("interiorFactor")
        return self._interior

    def condense_loads(self, loads):
        """f* = f_b − K_bi·K_ii⁻¹·f_i for each column of `loads` (dimension rows, in the matrix's numbering):
        trailing_dof_count × columns (PartialFactor::condenseLoads, PartialFactor.cpp:117-133)."""
        self._require_kept("condenseLoads")
        f = np.asfortranarray(self._columns("condenseLoads", loads, self._dimension, "the loads"), dtype=np.float64)
        t, k = self._trailing, f.shape[1]
        if t == 0:
            return np.zeros((0, k), order="F")
        out = np.zeros((t, k), order="F")
        _check(_native.lib().oosg_partial_condense_loads(self._interior.handle, _ptr(f), self._dimension, k,
                                                          _ptr(out), t))
        return out

    def recover(self, loads_to_solution, trailing_values):
        """Recovery (PartialFactor::recover, PartialFactor.cpp:135-158): the interior rows of `loads_to_solution`
        hold f_i (its trailing rows are ignored) and trailing_values holds u_b (trailing_dof_count × the same
        columns); returns u in every row (in place when `loads_to_solution` is a Fortran-ordered float64 array)."""
        self._require_kept("recover")
        x = self._columns("recover", loads_to_solution, self._dimension, "the loads")
        ub = self._columns("recover", trailing_values, self._trailing, "the trailing values")
        if ub.shape[1] != x.shape[1]:
            raise ValueError(f"PartialFactor::recover: {ub.shape[1]} columns of trailing values for {x.shape[1]} "
                             "load columns; give one per load column")
        inplace = isinstance(loads_to_solution, np.ndarray) and loads_to_solution.dtype == np.float64 and \
            loads_to_solution.flags.f_contiguous and loads_to_solution.ndim == 2
        x = x if inplace else np.array(x, dtype=np.float64, order="F", copy=True)
        if self._trailing == 0:
            x[:] = FactorSolve(self._interior).solve(x)
            return x
        ub = np.asfortranarray(ub, dtype=np.float64)
        _check(_native.lib().oosg_partial_recover(self._interior.handle, _ptr(x), self._dimension, x.shape[1],
                                                   _ptr(ub), self._trailing))
        return x

    def is_factorisation_of(self, symbolic, matrix, policy=None):
        """Whether the interior factor is kept and is the factorisation of `matrix` on `symbolic` with the pivot policy
        `policy` (default: this factor's): the shapes, the policy and the matrix checksum (PartialFactor.cpp:86-92)."""
        policy = policy or self._pivot_policy
        return (self._interior is not None and self._dimension == symbolic.dof_count == matrix.n and
                self._trailing == symbolic.trailing_dof_count and
                self._pivot_policy.max_ratio == policy.max_ratio and self._pivot_policy.bailout == policy.bailout and
                self._checksum == matrix_checksum(matrix))

    def release_interior_factor(self):
        """Frees the interior factor (K* stays): recovery needs recompute_interior_factor first."""
        self._interior = None

    def recompute_interior_factor(self, symbolic, matrix, options=None):
        """Computes the interior factor again from the same symbolic factor and matrix, with the same pivot policy:
        the same bits as the first time. ValueError when they do not have the shape of the first factorisation, when
        the matrix differs (its checksum) or the pivot policy differs (PartialFactor.cpp:111-150)."""
        options = options or FactorOptions()
        self._require_positive_definite(options)
        if symbolic.dof_count != self._dimension or symbolic.trailing_dof_count != self._trailing:
            raise ValueError(f"PartialFactor::recomputeInteriorFactor: the symbolic factor has {symbolic.dof_count} "
                             f"equations and {symbolic.trailing_dof_count} trailing ones, the partial factor "
                             f"{self._dimension} and {self._trailing}; pass the symbolic factor and the matrix of the "
                             "first factorisation")
        policy = options.pivot_policy
        if policy.max_ratio != self._pivot_policy.max_ratio or policy.bailout != self._pivot_policy.bailout:
            raise ValueError(f"PartialFactor::recomputeInteriorFactor: the pivot policy (ratio limit "
                             f"{policy.max_ratio}, bail-out {policy.bailout}) differs from the first factorisation's "
                             f"({self._pivot_policy.max_ratio}, {self._pivot_policy.bailout}); pass the options of "
                             "the first factorisation")
        if matrix_checksum(matrix) != self._checksum:
            raise ValueError("PartialFactor::recomputeInteriorFactor: the matrix differs from the one first factorised "
                             "(its checksum of structure and values changed); pass the same matrix, or factorise the "
                             "unit again")
        count = symbolic.supernode_count - (1 if self._trailing > 0 else 0)
        if count != self._interior_supernodes:
            raise ValueError("PartialFactor::recomputeInteriorFactor: the symbolic factor has another supernode "
                             "structure than the first factorisation; pass the same symbolic factor")
        self._interior = self._factorise_interior(symbolic, matrix, options)


def estimate_memory_bytes(symbolic, pivoting=0):
    """The device bytes oosg_plan_create reserves for `symbolic` (no device is touched)."""
    out = ctypes.c_uint64()
    desc = symbolic.desc()
    _check(_native.lib().oosg_estimate_memory_bytes(ctypes.byref(desc), int(pivoting), ctypes.byref(out)))
    return int(out.value)


def nested_dissection(graph_offsets, adjacency, node_weights, coordinates, leaf_nodes=64):
    """The node order of this project's geometric nested dissection (node at each position)."""
    lib = _native.lib()
    graph_offsets = np.ascontiguousarray(graph_offsets, dtype=np.int64)
    adjacency = np.ascontiguousarray(adjacency, dtype=np.int32)
    node_weights = np.ascontiguousarray(node_weights, dtype=np.int32)
    coordinates = np.ascontiguousarray(coordinates, dtype=np.float64).reshape(-1)
    n = int(graph_offsets.shape[0] - 1)
    order = np.zeros(n, np.int32)
    status = lib.oosa_order_dissection(n, _ptr(graph_offsets), _ptr(adjacency), _ptr(node_weights),
                                       _ptr(coordinates), int(leaf_nodes), _ptr(order))
    if status != 0:
        raise ValueError(lib.oosa_last_error().decode())
    return order
