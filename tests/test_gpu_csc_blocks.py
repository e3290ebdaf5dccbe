"""SymmetricCsc::fromBlocks on the GPU (oosg_csc_from_blocks_device) against the oracle's conversion, which is the
reference's bit for bit (tests/test_csc_blocks.py), and the hand-off of the device CSC to the factorisation.

Contract: offsets, rows and values bit-equal to the oracle's; a factor of the device-made matrix bit-equal to the
factor of the same matrix given from the host; malformed inputs refused (ValueError) before any term is written."""

import numpy as np
import pytest

import _oracle
from _problems import block_problem, grid_graph
from o_o_solver_amd import (BlockSparseMatrix, DeviceMatrix, DofSelection, SparsePattern, SupernodalFactor,
                            SymmetricCsc, csc_from_blocks_device, models)

pytestmark = pytest.mark.gpu


def _star(nodes):
    nb = [list(range(1, nodes))] + [[0] for _ in range(1, nodes)]
    offsets = np.concatenate([[0], np.cumsum([len(x) for x in nb])]).astype(np.int64)
    return offsets, np.array([v for x in nb for v in x], np.int32)


CASES = {
    "plate 20": lambda: (models.plate(20, jitter=0.2, seed=3).graph_offsets,
                         models.plate(20, jitter=0.2, seed=3).adjacency),
    "grid 30": lambda: grid_graph(30)[:2],
    "star 300": lambda: _star(300),  # a column of about a thousand rows: ranked from memory
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("eliminate", [0.0, 0.2])
def test_device_conversion_equals_the_oracle(gpu_device, case, eliminate):
    print(f"[gpu] fromBlocks {case} eliminate {eliminate}", flush=True)
    goff, adj = CASES[case]()
    counts, pattern, values, eliminated, order = block_problem(goff, adj, seed=7, eliminate=eliminate)
    sel = DofSelection(counts, eliminated, order)
    offsets, rows, vals = csc_from_blocks_device(BlockSparseMatrix(pattern, values), sel)
    ro, rr, rv = _oracle.oracle_csc_from_blocks(pattern, values, sel.equation_of_dof, sel.kept_dof_count)
    assert np.array_equal(offsets.cpu().numpy(), ro)
    assert np.array_equal(rows.cpu().numpy(), rr)
    assert np.array_equal(vals.cpu().numpy(), rv)


def test_device_matrix_hand_off_to_the_factorisation(gpu_device):
    """A plate's stiffness as node blocks (its CSC read back into blocks), converted on the device in the factor's
    order, handed to the plan without a host conversion: the factor is the host path's bit for bit, and new values
    set on the device give the host path's factor of the same values."""
    import torch
    w = models.plate(24, jitter=0.2, seed=5, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
    sym = w.analyse()
    m = w.matrix
    n = sym.n
    # Node blocks of the CSC (6 DOFs a node, the equations of node node_order[k] are 6k .. 6k + 5).
    counts = np.full(w.node_count, 6, np.uint8)
    pattern = SparsePattern.from_graph(w.graph_offsets, w.adjacency, counts)
    position = np.empty(w.node_count, np.int64)
    position[sym.node_order] = np.arange(w.node_count)
    dense_lower = {}
    for c in range(n):
        for k in range(m.offsets[c], m.offsets[c + 1]):
            dense_lower[(int(m.rows[k]), c)] = m.values[k]
    values = np.zeros(pattern.value_count)
    for j in range(w.node_count):
        for b in range(pattern.column_start[j], pattern.column_start[j + 1]):
            i = pattern.block_rows[b]
            block = np.zeros((6, 6))
            for cc in range(6):
                for rr in range(6):
                    p, q = 6 * position[i] + rr, 6 * position[j] + cc
                    block[rr, cc] = dense_lower.get((max(p, q), min(p, q)), 0.0)
            values[pattern.block_value_offsets[b]:pattern.block_value_offsets[b + 1]] = block.T.reshape(-1)
    sel = DofSelection(counts, None, sym.node_order)
    offsets, rows, vals = csc_from_blocks_device(BlockSparseMatrix(pattern, values), sel)
    # The pattern keeps structural zeros the model's CSC also keeps: the same structure here.
    assert np.array_equal(offsets.cpu().numpy(), m.offsets) and np.array_equal(rows.cpu().numpy(), m.rows)
    assert np.array_equal(vals.cpu().numpy(), m.values)
    plan = sym.plan()
    device_made = DeviceMatrix.from_device_csc(plan, offsets, rows, vals)
    f_dev = SupernodalFactor.factorise(sym, m, device_matrix=device_made)
    f_host = SupernodalFactor.factorise(sym, m)
    for s in range(0, sym.supernode_count, max(1, sym.supernode_count // 40)):
        assert np.array_equal(f_dev.panel(s)[0], f_host.panel(s)[0])
    # New values (the blocks doubled, converted on the device) set without a host copy.
    _, _, vals2 = csc_from_blocks_device(BlockSparseMatrix(pattern, 2.0 * values), sel)
    device_made.set_values_device(vals2)
    doubled = SymmetricCsc(n, m.offsets, m.rows, 2.0 * m.values)
    f_dev2 = SupernodalFactor.factorise(sym, doubled, device_matrix=device_made)
    f_host2 = SupernodalFactor.factorise(sym, doubled)
    assert np.array_equal(f_dev2.diagonal(), f_host2.diagonal())
    del f_dev, f_host, f_dev2, f_host2
    torch.cuda.synchronize()


def test_malformed_block_matrices_are_refused(gpu_device):
    goff, adj = grid_graph(8)[:2]
    counts, pattern, values, eliminated, order = block_problem(goff, adj, seed=3)
    sel = DofSelection(counts, eliminated, order)
    # Block rows out of order in a column (each stays a valid node).
    bad = SparsePattern(pattern.first_dof, pattern.column_start, pattern.block_rows.copy(),
                        pattern.block_value_offsets)
    j = next(j for j in range(pattern.node_count) if pattern.column_start[j + 1] - pattern.column_start[j] >= 3)
    b = pattern.column_start[j]
    bad.block_rows[b + 1], bad.block_rows[b + 2] = bad.block_rows[b + 2], bad.block_rows[b + 1]
    with pytest.raises(ValueError, match="not well formed"):
        csc_from_blocks_device(BlockSparseMatrix(bad, values), sel)
    # Two DOFs numbered as one equation.
    wrong = DofSelection(counts, eliminated, order)
    kept = np.flatnonzero(wrong.equation_of_dof >= 0)
    wrong.equation_of_dof[kept[1]] = wrong.equation_of_dof[kept[0]]
    with pytest.raises(ValueError, match="numbering of the kept DOFs"):
        csc_from_blocks_device(BlockSparseMatrix(pattern, values), wrong)
    # A selection of another pattern.
    other = DofSelection(np.append(counts, 1))
    with pytest.raises(ValueError, match="was not made for the pattern"):
        csc_from_blocks_device(BlockSparseMatrix(pattern, values), other)
