"""Synthetic finite-element workloads (liboosm.so): the matrices the factor and the solves are measured and tested
on. The reference assembles its stiffness matrices from its own elements (src/elements, src/assembly), which are
outside the numeric phase; these models are this project's own (flat MITC4 shells with a drilling penalty, trilinear
hexahedra, a stiffened cylinder with its geometric stiffness), generated from a seed so that a workload is the same
on every machine.

A Workload carries the node graph, coordinates and weights, the order (this project's nested dissection, or one
given), the symbolic factor and the lower CSC in factor order, and builds loads in that order.
"""

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native
from . import SymbolicFactor, SymmetricCsc, nested_dissection


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Workload:
    name: str
    handle: object
    node_count: int
    dof_per_node: int
    graph_offsets: np.ndarray
    adjacency: np.ndarray
    coordinates: np.ndarray
    weights: np.ndarray
    symbolic: SymbolicFactor = None
    matrix: SymmetricCsc = None

    def __del__(self):
        handle = getattr(self, "handle", None)
        self.handle = None
        if handle and _native is not None and getattr(_native, "models", None) is not None:
            try:
                _native.models().oosm_free(handle)
            except (AttributeError, TypeError, ImportError):
                pass  # interpreter shutdown: the library's module state is already gone

    @property
    def n(self):
        return self.node_count * self.dof_per_node

    def analyse(self, order=None, amalgamate=True):
        """The symbolic factor under `order` (default: nested dissection) and the CSC in its order."""
        if order is None:
            order = nested_dissection(self.graph_offsets, self.adjacency, self.weights, self.coordinates)
        self.symbolic = SymbolicFactor.analyse(self.graph_offsets, self.adjacency, self.weights, order,
                                               amalgamate=amalgamate)
        self.matrix = self.csc(self.symbolic.node_order)
        return self.symbolic

    def csc(self, node_order):
        m = _native.models()
        sizes = np.zeros(4, np.int64)
        m.oosm_sizes(self.handle, _ptr(sizes))
        nnz = int(sizes[3])
        offsets = np.zeros(self.n + 1, np.int64)
        rows = np.zeros(nnz, np.int32)
        values = np.zeros(nnz)
        node_order = np.ascontiguousarray(node_order, dtype=np.int32)
        m.oosm_csc(self.handle, _ptr(node_order), _ptr(offsets), _ptr(rows), _ptr(values))
        return SymmetricCsc(self.n, offsets, rows, values)

    def geometric(self):
        """The geometric stiffness K_G on the stiffness's pattern, in the factor's order (zero for models without
        one); the buckling factors λ solve (K + λ·K_G)φ = 0."""
        m = _native.models()
        nnz = self.matrix.rows.shape[0]
        offsets = np.zeros(self.n + 1, np.int64)
        rows = np.zeros(nnz, np.int32)
        values = np.zeros(nnz)
        node_order = np.ascontiguousarray(self.symbolic.node_order, dtype=np.int32)
        m.oosm_csc_geometric(self.handle, _ptr(node_order), _ptr(offsets), _ptr(rows), _ptr(values))
        return SymmetricCsc(self.n, offsets, rows, values)

    def loads(self, kind="random", nrhs=1, seed=1):
        """n × nrhs (Fortran order) in the factor's equation order. kind: 'pressure' (uniform pressure, shells),
        'random' (N(0, 1) forces on the translational DOFs), 'top' (N(0, 1) on the top face's nodes)."""
        code = {"pressure": 0, "random": 1, "top": 2}[kind]
        b = np.zeros((nrhs, self.n))
        _native.models().oosm_loads(self.handle, _ptr(self.symbolic.node_order), code, int(nrhs), int(seed), _ptr(b))
        return np.asfortranarray(b.T)


def _workload(name, handle):
    m = _native.models()
    sizes = np.zeros(4, np.int64)
    m.oosm_sizes(handle, _ptr(sizes))
    n_nodes, dof = int(sizes[0]), int(sizes[1])
    offsets = np.zeros(n_nodes + 1, np.int64)
    adjacency = np.zeros(int(sizes[2]), np.int32)
    m.oosm_graph(handle, _ptr(offsets), _ptr(adjacency))
    coords = np.zeros(3 * n_nodes)
    m.oosm_coords(handle, _ptr(coords))
    return Workload(name, handle, n_nodes, dof, offsets, adjacency, coords, np.full(n_nodes, dof, np.int32))


def plate(nx, ny=None, length=1.0, width=None, jitter=0.0, seed=0, thickness=(2e-3, 2e-3), thickness_mode="constant",
          E=70e9, nu=0.3, pressure=1e3):
    """A flat plate of nx × ny quadrilateral shells, all edges clamped (6 DOFs per kept node)."""
    ny = nx if ny is None else ny
    width = length if width is None else width
    mode = {"constant": 0, "log-uniform": 1, "uniform": 2}[thickness_mode]
    handle = _native.models().oosm_plate(nx, ny, length, width, jitter, seed, thickness[0], thickness[1], mode, E, nu,
                                         pressure)
    return _workload(f"plate {nx}x{ny}", handle)


def hex_block(nx, ny=None, nz=None, size=(1.0, 1.0, 1.0), seed=0, E_median=200e9, sigma=0.0, nu=0.3):
    """A block of nx × ny × nz hexahedra, the face z = 0 fixed (3 DOFs per kept node)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    handle = _native.models().oosm_hex_block(nx, ny, nz, size[0], size[1], size[2], seed, E_median, sigma, nu)
    return _workload(f"hex block {nx}x{ny}x{nz}", handle)


def cylinder(n_circ=360, n_axial=200, radius=1.0, length=2.0, thickness=2e-3, E=70e9, nu=0.3, stringer_every=15,
             stringer_area=2e-4, axial_force=-1e3):
    """A stiffened cylinder (the buckling workload): n_circ × n_axial quad4 shells on a cylinder of radius `radius`
    and length `length`, both end rings fixed, bar stringers along every `stringer_every`-th meridian, and the
    geometric stiffness of the uniform axial membrane force `axial_force` (N/m, compression < 0)."""
    handle = _native.models().oosm_cylinder(n_circ, n_axial, radius, length, thickness, E, nu, stringer_every,
                                            stringer_area, axial_force)
    return _workload(f"cylinder {n_circ}x{n_axial}", handle)


def shell_plate_c2(scale=500, seed=20261203):
    """The shell plate of the headline workload: scale × scale quad4 over 1 m, interior nodes jittered ±20 % of the
    element size, per-element thickness log-uniform in [1, 4] mm, E = 70 GPa, ν = 0.3, clamped edges."""
    return plate(scale, jitter=0.2, seed=seed, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")


class Batch:
    """Independent models factorised as one: the block-diagonal union of their matrices, whose elimination forest
    has one tree per model, so that every level's fronts of all the models go into the same launches. Each model is
    ordered by its own dissection; its equations are one contiguous range of the union's."""

    def __init__(self, members, name=None):
        self.members = list(members)
        self.name = name or f"batch of {len(self.members)}"
        self.symbolic = None
        self.matrix = None
        self.ranges = None  # per member (first equation, equation count), in factor order

    @property
    def n(self):
        return sum(w.n for w in self.members)

    def analyse(self):
        node_base = np.cumsum([0] + [w.node_count for w in self.members])
        offsets, adjacency, weights, order = [np.zeros(1, np.int64)], [], [], []
        edge_base = 0
        for w, base in zip(self.members, node_base[:-1]):
            offsets.append(w.graph_offsets[1:] + edge_base)
            edge_base += int(w.graph_offsets[-1])
            adjacency.append(w.adjacency + base)
            weights.append(w.weights)
            order.append(nested_dissection(w.graph_offsets, w.adjacency, w.weights, w.coordinates) + base)
        self.graph_offsets = np.concatenate(offsets)
        self.adjacency = np.concatenate(adjacency).astype(np.int32)
        self.weights = np.concatenate(weights).astype(np.int32)
        self.symbolic = SymbolicFactor.analyse(self.graph_offsets, self.adjacency, self.weights,
                                               np.concatenate(order).astype(np.int32))
        # Each model is one tree of the forest: its nodes are one contiguous run of the final order.
        node_order = self.symbolic.node_order
        owner = np.searchsorted(node_base, node_order, side="right") - 1
        starts = np.flatnonzero(np.r_[True, owner[1:] != owner[:-1]])
        if len(starts) != len(self.members):
            raise ValueError("batch: a model's nodes are not contiguous in the factor order")
        dof_of_position = np.concatenate([[0], np.cumsum(self.weights[node_order])])
        self.ranges = [None] * len(self.members)
        columns, rows, values, col_offsets = [], [], [], [np.zeros(1, np.int64)]
        nnz = 0
        for q, start in enumerate(starts):
            m = int(owner[start])
            w = self.members[m]
            local_order = (node_order[start:start + w.node_count] - node_base[m]).astype(np.int32)
            w.symbolic = None
            csc = w.csc(local_order)
            w.local_order = local_order
            eq0 = int(dof_of_position[start])
            self.ranges[m] = (eq0, w.n)
            col_offsets.append(csc.offsets[1:] + nnz)
            nnz += int(csc.offsets[-1])
            rows.append(csc.rows + eq0)
            values.append(csc.values)
        self.matrix = SymmetricCsc(self.n, np.concatenate(col_offsets), np.concatenate(rows).astype(np.int32),
                                   np.concatenate(values))
        return self.symbolic

    def loads(self, kind="random", nrhs=1, seed=1):
        """The members' loads (member m with seed + m) stacked in the union's equation order."""
        b = np.zeros((self.n, nrhs), order="F")
        for m, w in enumerate(self.members):
            eq0, count = self.ranges[m]
            block = np.zeros((nrhs, w.n))
            _native.models().oosm_loads(w.handle, _ptr(w.local_order), {"pressure": 0, "random": 1, "top": 2}[kind],
                                        int(nrhs), int(seed + m), _ptr(block))
            b[eq0:eq0 + count, :] = block.T
        return b


def plate_batch_c4(count=512, seed=20261204, sides=(20, 80)):
    """The batch of clamped square plates: side n_e ~ U{20..80} elements over 1 m, thickness ~ U[1, 5] mm per
    element, E = 70 GPa, ν = 0.3."""
    rng = np.random.default_rng(seed)
    members = [plate(int(rng.integers(sides[0], sides[1] + 1)), seed=seed + q, thickness=(1e-3, 5e-3),
                     thickness_mode="uniform") for q in range(count)]
    return Batch(members, name=f"batch of {count} plates")
