"""Small scalar test problems in the form of the reference's checks (checks/linalg/FactorEdgeChecks.cpp:50-104,
checks/linalg/FactorOrderChecks.cpp:54-114): node graphs of chains and grids, and their matrices assembled in the
equation order of a symbolic factor."""

import numpy as np

from o_o_solver_amd import SymmetricCsc


def chain_graph(count):
    """A chain of `count` scalar nodes (node v next to v − 1 and v + 1)."""
    offsets, nb = [0], []
    for v in range(count):
        if v > 0:
            nb.append(v - 1)
        if v + 1 < count:
            nb.append(v + 1)
        offsets.append(len(nb))
    return (np.array(offsets, np.int64), np.array(nb, np.int32), np.ones(count, np.int32),
            np.stack([np.arange(count, dtype=float), np.zeros(count), np.zeros(count)], 1))


def grid_graph(side):
    """side × side scalar nodes, each next to its four neighbours (node i + side·j)."""
    offsets, nb = [0], []
    for j in range(side):
        for i in range(side):
            if j > 0:
                nb.append(i + side * (j - 1))
            if i > 0:
                nb.append(i - 1 + side * j)
            if i + 1 < side:
                nb.append(i + 1 + side * j)
            if j + 1 < side:
                nb.append(i + side * (j + 1))
            offsets.append(len(nb))
    coords = np.array([[i, j, 0.0] for j in range(side) for i in range(side)], float)
    return np.array(offsets, np.int64), np.array(nb, np.int32), np.ones(side * side, np.int32), coords


def graph_matrix(symbolic, graph_offsets, adjacency, diagonal, coupling):
    """The matrix of a scalar graph in the equations of `symbolic`: diagonal[v] at node v, coupling(a, b) between
    neighbours a < b; each column's rows ascending."""
    position = np.empty(len(symbolic.node_order), np.int64)
    position[symbolic.node_order] = np.arange(len(symbolic.node_order))
    n = len(position)
    columns = [[] for _ in range(n)]
    for a in range(n):
        ea = position[a]
        columns[ea].append((ea, float(diagonal[a])))
        for b in adjacency[graph_offsets[a]:graph_offsets[a + 1]]:
            if b > a:
                eb = position[b]
                columns[min(ea, eb)].append((max(ea, eb), float(coupling(a, b))))
    offsets, rows, values = [0], [], []
    for col in columns:
        col.sort()
        rows.extend(r for r, _ in col)
        values.extend(v for _, v in col)
        offsets.append(len(rows))
    return SymmetricCsc(n, np.array(offsets, np.int64), np.array(rows, np.int32), np.array(values))


def random_spd(n, seed, shift=None):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    a = a @ a.T
    a += np.eye(n) * (n if shift is None else shift)
    return a


def nested_chain_order(count):
    """A nested-dissection order of a chain: the middle node last, the halves ordered the same way before it."""
    order = []

    def visit(lo, hi):
        if lo >= hi:
            return
        mid = (lo + hi) // 2
        visit(lo, mid)
        visit(mid + 1, hi)
        order.append(mid)

    visit(0, count)
    return np.array(order, np.int32)


def partial_plate(nx, jitter=0.0, seed=0, thickness_mode="constant"):
    """A clamped plate cut by the line of nodes nearest x = 0.4 (its unit's boundary, the trailing set, ordered last
    in the order of their node numbers), the rest in nested dissection order: (workload, symbolic factor with its
    trailing set, matrix in its order, the trailing node count)."""
    from o_o_solver_amd import SymbolicFactor, models, nested_dissection
    w = models.plate(nx, jitter=jitter, seed=seed, thickness=(1e-3, 5e-3) if thickness_mode != "constant" else
                     (2e-3, 2e-3), thickness_mode=thickness_mode)
    x = np.asarray(w.coordinates).reshape(-1, 3)[:, 0]
    line = np.abs(x - 0.4)
    trailing = np.flatnonzero(line <= line.min() + 0.5 / nx)
    order = nested_dissection(w.graph_offsets, w.adjacency, w.weights, w.coordinates)
    keep = np.ones(w.node_count, bool)
    keep[trailing] = False
    order = np.concatenate([order[keep[order]], trailing]).astype(np.int32)
    sym = SymbolicFactor.analyse(w.graph_offsets, w.adjacency, w.weights, order, trailing_node_count=len(trailing))
    w.symbolic = sym
    w.matrix = w.csc(sym.node_order)
    return w, sym, w.matrix, len(trailing), order


def block_problem(graph_offsets, adjacency, seed, max_dofs=6, eliminate=0.1, ordered=True):
    """A block matrix on the pattern of a node graph: random DOF counts (1 to max_dofs per node), random values (the
    diagonal blocks symmetric), a random fraction of DOFs eliminated (each node keeps one at least when ordered) and
    a random node order: (dof_counts, pattern, values, eliminated flags, node order)."""
    from o_o_solver_amd import SparsePattern
    rng = np.random.default_rng(seed)
    nodes = len(graph_offsets) - 1
    counts = rng.integers(1, max_dofs + 1, nodes).astype(np.uint8)
    pattern = SparsePattern.from_graph(graph_offsets, adjacency, counts)
    values = rng.standard_normal(pattern.value_count)
    # A diagonal block is stored complete (both triangles): make it symmetric.
    for j in range(nodes):
        b = pattern.column_start[j]
        d = int(counts[j])
        v = values[pattern.block_value_offsets[b]:pattern.block_value_offsets[b + 1]].reshape(d, d)
        v[:] = (v + v.T) / 2
    eliminated = (rng.random(pattern.dof_count) < eliminate).astype(np.uint8)
    order = rng.permutation(nodes).astype(np.int32) if ordered else None
    return counts, pattern, values, eliminated, order
