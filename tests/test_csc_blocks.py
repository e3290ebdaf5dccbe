"""The CSC of a block matrix (SymmetricCsc::fromBlocks, src/linalg/sparse/SymmetricCsc.cpp:41-200): the host mirror's
pattern and selection and the oracle's conversion against the compiled reference, bit for bit."""

import numpy as np
import pytest

import _oracle
from _problems import block_problem, chain_graph, grid_graph
from o_o_solver_amd import DofSelection, SparsePattern, models


def _graphs():
    w = models.plate(6, jitter=0.1, seed=2)
    yield "plate 6", w.graph_offsets, w.adjacency
    g = grid_graph(9)
    yield "grid 9", g[0], g[1]
    g = chain_graph(40)
    yield "chain 40", g[0], g[1]
    # A star: node 0 next to every other (a column of several hundred rows).
    nodes = 120
    nb = [list(range(1, nodes))] + [[0] for _ in range(1, nodes)]
    offsets = np.concatenate([[0], np.cumsum([len(x) for x in nb])]).astype(np.int64)
    yield "star 120", offsets, np.array([v for x in nb for v in x], np.int32)


GRAPHS = list(_graphs())


@pytest.mark.parametrize("case", range(len(GRAPHS)), ids=[g[0] for g in GRAPHS])
def test_pattern_and_selection_equal_the_reference(reference_lib, case):
    name, goff, adj = GRAPHS[case]
    counts, pattern, values, eliminated, order = block_problem(goff, adj, seed=case)
    first, cols, rows, voff = _oracle.ref_pattern(goff, adj, counts)
    assert np.array_equal(pattern.first_dof, first)
    assert np.array_equal(pattern.column_start, cols)
    assert np.array_equal(pattern.block_rows, rows)
    assert np.array_equal(pattern.block_value_offsets, voff)
    for elim, node_order in [(eliminated, order), (None, None), (eliminated, None)]:
        sel = DofSelection(counts, elim, node_order)
        eq, kept = _oracle.ref_selection(counts, elim, node_order)
        assert np.array_equal(sel.equation_of_dof, eq) and sel.kept_dof_count == kept


@pytest.mark.parametrize("case", range(len(GRAPHS)), ids=[g[0] for g in GRAPHS])
@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_conversion_equals_the_reference(reference_lib, case, seed):
    name, goff, adj = GRAPHS[case]
    counts, pattern, values, eliminated, order = block_problem(goff, adj, seed=10 * case + seed)
    sel = DofSelection(counts, eliminated, order)
    mine = _oracle.oracle_csc_from_blocks(pattern, values, sel.equation_of_dof, sel.kept_dof_count)
    ref = _oracle.ref_from_blocks(goff, adj, counts, values, eliminated, order, threads=2)
    for a, b in zip(mine, ref):
        assert np.array_equal(a, b)
    # The dense matrix of the blocks, kept DOFs in equation order, is the CSC's.
    n = pattern.dof_count
    dense = np.zeros((n, n))
    for j in range(pattern.node_count):
        for b in range(pattern.column_start[j], pattern.column_start[j + 1]):
            i = pattern.block_rows[b]
            di, dj = pattern.first_dof[i + 1] - pattern.first_dof[i], pattern.first_dof[j + 1] - pattern.first_dof[j]
            block = values[pattern.block_value_offsets[b]:pattern.block_value_offsets[b + 1]].reshape(dj, di).T
            dense[pattern.first_dof[i]:pattern.first_dof[i] + di, pattern.first_dof[j]:pattern.first_dof[j] + dj] = block
    dense = np.tril(dense) + np.tril(dense, -1).T
    keep = np.flatnonzero(sel.equation_of_dof >= 0)
    perm = np.empty(sel.kept_dof_count, np.int64)
    perm[sel.equation_of_dof[keep]] = keep
    expected = dense[np.ix_(perm, perm)]
    offsets, rows, vals = mine
    got = np.zeros_like(expected)
    for c in range(sel.kept_dof_count):
        got[rows[offsets[c]:offsets[c + 1]], c] = vals[offsets[c]:offsets[c + 1]]
    assert np.array_equal(np.tril(expected), got)
