"""Writes the golden fixtures of tests/golden from the compiled reference (oracle/_ref/liboos_ref.so, built from
/root/reference by oracle/build_ref.py): inputs and the reference's outputs, so that the GPU tests can check against
the reference on a machine where it is not present. The reference has no stored fixtures of its own for this path.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  plate8.npz     8 × 8 clamped shell plate (jitter 0.2, thickness log-uniform [1, 4] mm, seed 11), the order of this
                 project's nested dissection; the symbolic factor, the CSC in factor order, two N(0, 1) load columns,
                 the reference's solution (reference kernel family), D per equation and the inertia.
  c1plate24.npz  24 × 24 clamped shell plate of the statics configuration's kind (1 m square, t = 2 mm constant, no
                 jitter, E = 70 GPa, ν = 0.3, uniform pressure 1 kPa as consistent nodal forces, one load case), the
                 order of this project's nested dissection: the symbolic factor, the CSC, the load, the reference's
                 solution, D and inertia, and the equation of the centre node's deflection w.
  grid20.npz     20 × 20 scalar grid, −1 couplings, the reference's minimum-degree order; diagonals 4.5 (positive
                 definite) and 3.3 (indefinite): the reference's inertia of each (positive definite and Bunch-Kaufman
                 factorisations) and its solutions of the right-hand sides (7i + 3j) mod 11 − 5
                 (checks/linalg/FactorOrderChecks.cpp:127-129).
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _oracle  # noqa: E402
from _problems import graph_matrix, grid_graph  # noqa: E402
from o_o_solver_amd import SymbolicFactor, models  # noqa: E402


def symbolic_arrays(sym):
    return dict(node_order=sym.node_order, first_column=sym.first_column, column_count=sym.column_count,
                parent=sym.parent, row_start=sym.row_start, rows=sym.rows_all)


def plate8():
    w = models.plate(8, jitter=0.2, seed=11, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
    sym = w.analyse()
    m = w.matrix
    b = w.loads("random", nrhs=2, seed=5)
    x = b.copy(order="F")
    h, sizes, flops, arr = _oracle.ref_symbolic(w.graph_offsets, w.adjacency, w.weights, order=sym.node_order.copy())
    try:
        rc, msg, inertia, counts, times, panels, diag = _oracle.ref_factor_solve(h, m, b=x, kernel=0)
    finally:
        _oracle.reference().ref_sym_free(h)
    assert rc == 0, msg
    np.savez_compressed(os.path.join(HERE, "plate8.npz"), n=m.n, offsets=m.offsets, csc_rows=m.rows,
                        values=m.values, b=b, x=x, diag=diag, inertia=inertia, flops=flops, **symbolic_arrays(sym))


def centre_equation(w, sym):
    """The equation of the w DOF (the third of six) of the node nearest the plate's centre."""
    xy = w.coordinates.reshape(-1, 3)[:, :2]
    node = int(np.argmin(np.sum((xy - xy.mean(axis=0)) ** 2, axis=1)))
    position = int(np.flatnonzero(sym.node_order == node)[0])
    return position * w.dof_per_node + 2


def c1plate24():
    w = models.plate(24, thickness=(2e-3, 2e-3), thickness_mode="constant", pressure=1e3)
    sym = w.analyse()
    m = w.matrix
    b = w.loads("pressure", nrhs=1)
    x = b.copy(order="F")
    h, sizes, flops, arr = _oracle.ref_symbolic(w.graph_offsets, w.adjacency, w.weights, order=sym.node_order.copy())
    try:
        rc, msg, inertia, counts, times, panels, diag = _oracle.ref_factor_solve(h, m, b=x, kernel=0)
    finally:
        _oracle.reference().ref_sym_free(h)
    assert rc == 0, msg
    np.savez_compressed(os.path.join(HERE, "c1plate24.npz"), n=m.n, offsets=m.offsets, csc_rows=m.rows,
                        values=m.values, b=b, x=x, diag=diag, inertia=inertia, flops=flops,
                        centre=centre_equation(w, sym), **symbolic_arrays(sym))


def grid20():
    g = grid_graph(20)
    h, sizes, flops, arr = _oracle.ref_symbolic(g[0], g[1], g[2], method=1)
    _oracle.reference().ref_sym_free(h)
    sym = SymbolicFactor.analyse(g[0], g[1], g[2], arr["node_order"])
    out = dict(n=400, **symbolic_arrays(sym))
    b = np.array([[(7 * i + 3 * j) % 11 - 5.0 for j in range(2)] for i in range(400)], order="F")
    out["b"] = b
    for name, diagonal, pivoting in (("pd", 4.5, 0), ("indefinite", 3.3, 1)):
        m = graph_matrix(sym, g[0], g[1], np.full(400, diagonal), lambda a, c: -1.0)
        h, *_ = _oracle.ref_symbolic(g[0], g[1], g[2], order=sym.node_order.copy())
        x = b.copy(order="F")
        try:
            rc, msg, inertia, counts, times, panels, diag = _oracle.ref_factor_solve(h, m, b=x, kernel=0,
                                                                                    pivoting=pivoting)
        finally:
            _oracle.reference().ref_sym_free(h)
        assert rc == 0, msg
        out[f"{name}_offsets"] = m.offsets
        out[f"{name}_csc_rows"] = m.rows
        out[f"{name}_values"] = m.values
        out[f"{name}_x"] = x
        out[f"{name}_inertia"] = inertia
        out[f"{name}_counts"] = counts
    np.savez_compressed(os.path.join(HERE, "grid20.npz"), **out)


if __name__ == "__main__":
    if _oracle.reference() is None:
        sys.exit("the compiled reference is not built: python oracle/build_ref.py")
    plate8()
    c1plate24()
    grid20()
    for name in sorted(os.listdir(HERE)):
        if name.endswith(".npz"):
            print(name, os.path.getsize(os.path.join(HERE, name)), "bytes")
