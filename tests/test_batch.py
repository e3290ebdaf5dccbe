"""Batches of independent models factorised as one (the block-diagonal union, one elimination tree per model): the
union's structure and factor are the members' own."""

import numpy as np

import _oracle
from o_o_solver_amd import SymbolicFactor, models


def _batch():
    bt = models.plate_batch_c4(count=5, seed=11, sides=(4, 9))
    bt.analyse()
    return bt


def test_batch_structure_is_the_members_structures():
    bt = _batch()
    sym = bt.symbolic
    supernodes, flops = 0, 0.0
    for w in bt.members:
        own = SymbolicFactor.analyse(w.graph_offsets, w.adjacency, w.weights, w.local_order)
        assert np.array_equal(own.node_order, w.local_order)
        supernodes += own.supernode_count
        flops += own.flop_count
    assert sym.supernode_count == supernodes
    assert abs(sym.flop_count - flops) <= 1e-9 * flops
    # One root per model.
    assert int(np.sum(sym.parent < 0)) == len(bt.members)


def test_batch_oracle_factor_is_the_members_factors():
    bt = _batch()
    _, diag, res = _oracle.multifrontal(bt.symbolic, bt.matrix)
    assert res.status == 0 and tuple(res.inertia) == (bt.n, 0, 0)
    for m, w in enumerate(bt.members):
        own = SymbolicFactor.analyse(w.graph_offsets, w.adjacency, w.weights, w.local_order)
        _, d_own, _ = _oracle.multifrontal(own, w.csc(w.local_order))
        eq0, count = bt.ranges[m]
        assert np.array_equal(diag[eq0:eq0 + count], d_own)
