"""The host symbolic analysis (liboosg.so) gives the reference's structure, bit for bit, for the same graph, weights
and order (SymbolicFactor::analyse, src/linalg/symbolic/SymbolicFactor.cpp:166-296)."""

import numpy as np
import pytest

import _oracle
from _problems import chain_graph, grid_graph
from o_o_solver_amd import SymbolicFactor, models, nested_dissection


def _compare(graph_offsets, adjacency, weights, order, amalgamate=True):
    mine = SymbolicFactor.analyse(graph_offsets, adjacency, weights, order, amalgamate=amalgamate)
    h, sizes, flops, arr = _oracle.ref_symbolic(graph_offsets, adjacency, weights, order=np.array(order, np.int32),
                                                amalgamate=1 if amalgamate else 0)
    _oracle.reference().ref_sym_free(h)
    assert np.array_equal(mine.node_order, arr["node_order"])
    assert np.array_equal(mine.first_column, arr["first_column"])
    assert np.array_equal(mine.column_count, arr["column_count"])
    assert np.array_equal(mine.parent, arr["parent"])
    assert np.array_equal(mine.row_start, arr["row_start"])
    assert np.array_equal(mine.rows_all, arr["rows"])
    assert mine.flop_count == flops
    assert mine.panel_value_count == sizes[4] and mine.factor_entry_count == sizes[5]
    return mine


@pytest.mark.parametrize("amalgamate", [True, False])
def test_chain(reference_lib, amalgamate):
    g = chain_graph(3000)
    _compare(g[0], g[1], g[2], np.arange(3000, dtype=np.int32), amalgamate)


@pytest.mark.parametrize("amalgamate", [True, False])
def test_grid_minimum_degree_order(reference_lib, amalgamate):
    g = grid_graph(60)
    h, sizes, flops, arr = _oracle.ref_symbolic(g[0], g[1], g[2], method=1)
    reference_lib.ref_sym_free(h)
    _compare(g[0], g[1], g[2], arr["node_order"], amalgamate)


def test_grid_random_order(reference_lib):
    g = grid_graph(40)
    order = np.random.default_rng(4).permutation(1600).astype(np.int32)
    _compare(g[0], g[1], g[2], order)


@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_plate_dissection(reference_lib, jitter):
    w = models.plate(40, jitter=jitter, seed=7)
    order = nested_dissection(w.graph_offsets, w.adjacency, w.weights, w.coordinates)
    sym = _compare(w.graph_offsets, w.adjacency, w.weights, order)
    # The dissection is a permutation whose flops stay near the reference's own dissection of the same plate.
    assert np.array_equal(np.sort(order), np.arange(w.node_count))
    h, sizes, flops, arr = _oracle.ref_symbolic(w.graph_offsets, w.adjacency, w.weights, coords=w.coordinates,
                                                method=2)
    reference_lib.ref_sym_free(h)
    assert sym.flop_count < 1.25 * flops


def test_hex_block_dissection(reference_lib):
    w = models.hex_block(8)
    order = nested_dissection(w.graph_offsets, w.adjacency, w.weights, w.coordinates)
    _compare(w.graph_offsets, w.adjacency, w.weights, order)


def test_invariants_without_reference():
    # Invariants that need no reference: parents follow children, child rows lie in the parent's columns and rows,
    # panel sizes and the flop formula.
    w = models.plate(20, jitter=0.2, seed=3)
    sym = w.analyse()
    S = sym.supernode_count
    assert sym.first_column[0] == 0 and sym.first_column[-1] + sym.column_count[-1] == sym.n
    for s in range(S):
        p = sym.parent[s]
        rows = sym.rows(s)
        assert np.all(np.diff(rows) > 0)
        if p < 0:
            assert len(rows) == 0
            continue
        assert p > s
        parent_set = set(range(sym.first_column[p], sym.first_column[p] + sym.column_count[p])) | set(
            sym.rows(p).tolist())
        assert set(rows.tolist()) <= parent_set
    c = sym.column_count.astype(np.int64)
    r = np.diff(sym.row_start)
    assert sym.panel_value_count == int(np.sum(c * (c + r)))
    flops = sum(float(m * (m + 1) + 2 * m) for cc, rr in zip(c, r) for m in range(rr, rr + cc))
    assert sym.flop_count == flops


def test_invalid_order_is_refused():
    g = chain_graph(5)
    with pytest.raises(ValueError):
        SymbolicFactor.analyse(g[0], g[1], g[2], np.array([0, 1, 2, 2, 4], np.int32))
