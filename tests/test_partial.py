"""The partial factorisation's checker (PartialFactor, src/linalg/supernodal/PartialFactor.cpp:65-158) against the
compiled reference and the dense closed forms, and the host side of the GPU mirror (no device calls)."""

import numpy as np
import pytest

import _oracle
from _problems import partial_plate
from o_o_solver_amd import (FactorOptions, PartialFactor, PivotingMode, SymbolicFactor, SymmetricCsc,
                            matrix_checksum)


def _dense_parts(matrix, interior):
    a = matrix.to_dense()
    return a[:interior, :interior], a[:interior, interior:], a[interior:, interior:]


@pytest.mark.parametrize("nx,jitter", [(12, 0.2), (20, 0.0)])
def test_trailing_structure_equals_the_reference(reference_lib, nx, jitter):
    w, sym, m, t, order = partial_plate(nx, jitter=jitter, seed=3)
    h, sizes, flops, arr = _oracle.ref_symbolic_trailing(w.graph_offsets, w.adjacency, w.weights, order, t)
    reference_lib.ref_sym_free(h)
    assert np.array_equal(sym.node_order, arr["node_order"])
    assert np.array_equal(sym.first_column, arr["first_column"])
    assert np.array_equal(sym.column_count, arr["column_count"])
    assert np.array_equal(sym.parent, arr["parent"])
    assert np.array_equal(sym.row_start, arr["row_start"])
    assert np.array_equal(sym.rows_all, arr["rows"])
    assert sym.flop_count == flops
    # The trailing set is the last supernode, with no rows below it.
    assert sym.trailing_dof_count == 6 * t == sym.column_count[-1]
    assert sym.row_count(sym.supernode_count - 1) == 0


@pytest.mark.parametrize("nx,jitter", [(12, 0.2), (20, 0.0)])
def test_oracle_partial_factor_equals_the_reference(reference_lib, nx, jitter):
    """K*, the interior panels and the inertia bit for bit (reference family); the condensation and the recovery to
    rounding (the sweeps' chains differ), and all of them against the dense formulas."""
    w, sym, m, t, order = partial_plate(nx, jitter=jitter, seed=3)
    h, _, _, _ = _oracle.ref_symbolic_trailing(w.graph_offsets, w.adjacency, w.weights, order, t)
    try:
        panels, diag, schur, res = _oracle.multifrontal_partial(sym, m)
        assert res.status == 0
        interior = sym.n - sym.trailing_dof_count
        values = sym.panel_value_count - int(sym.column_count[-1]) ** 2
        rng = np.random.default_rng(nx)
        f = rng.standard_normal((sym.n, 3))
        ub = rng.standard_normal((sym.trailing_dof_count, 3))
        ks, kp, kd, inertia, condensed, recovered = _oracle.ref_partial(h, m, sym.trailing_dof_count, values, f, ub)
    finally:
        reference_lib.ref_sym_free(h)
    assert np.array_equal(schur, ks)
    assert np.array_equal(panels[:values], kp)
    assert np.array_equal(diag[:interior], kd[:interior])
    assert list(res.inertia) == list(inertia) == [interior, 0, 0]
    co = _oracle.oracle_condense(sym, panels, f)
    ro = _oracle.oracle_recover(sym, panels, f, ub)
    assert np.max(np.abs(co - condensed)) <= 1e-13 * np.max(np.abs(condensed))
    assert np.max(np.abs(ro - recovered)) <= 1e-13 * np.max(np.abs(recovered))
    assert np.array_equal(ro[interior:], ub)
    # The formulas: K* = K_bb − K_bi·K_ii⁻¹·K_ib, f* = f_b − K_bi·K_ii⁻¹·f_i, u_i = K_ii⁻¹·(f_i − K_ib·u_b).
    kii, kib, kbb = _dense_parts(m, interior)
    dense = kbb - kib.T @ np.linalg.solve(kii, kib)
    assert np.max(np.abs(schur - dense)) <= 1e-12 * np.max(np.abs(dense))
    assert np.array_equal(schur, schur.T)
    fd = f[interior:] - kib.T @ np.linalg.solve(kii, f[:interior])
    assert np.max(np.abs(co - fd)) <= 1e-9 * np.max(np.abs(fd))
    ud = np.linalg.solve(kii, f[:interior] - kib @ ub)
    assert np.max(np.abs(ro[:interior] - ud)) <= 1e-9 * np.max(np.abs(ud))


def test_condensed_system_gives_the_whole_solution():
    """u_b = K*⁻¹·f*, recovered, solves K·u = f (the condensation is exact): the oracle against the whole solve."""
    w, sym, m, t, order = partial_plate(16, jitter=0.1, seed=5)
    panels, diag, schur, res = _oracle.multifrontal_partial(sym, m)
    f = np.random.default_rng(2).standard_normal((sym.n, 2))
    ub = np.linalg.solve(schur, _oracle.oracle_condense(sym, panels, f))
    u = _oracle.oracle_recover(sym, panels, f, ub)
    assert _oracle.backward_error(m, u, f) < 1e-14


def test_checksum_tracks_structure_and_values():
    w, sym, m, t, order = partial_plate(8, seed=1)
    same = SymmetricCsc(m.n, m.offsets.copy(), m.rows.copy(), m.values.copy())
    assert matrix_checksum(m) == matrix_checksum(same)
    changed = SymmetricCsc(m.n, m.offsets.copy(), m.rows.copy(), m.values.copy())
    changed.values[0] = np.nextafter(changed.values[0], np.inf)  # one bit of the first diagonal
    assert matrix_checksum(changed) != matrix_checksum(m)


def test_partial_factor_refuses_indefinite_mode_before_any_work():
    w, sym, m, t, order = partial_plate(8, seed=1)
    with pytest.raises(ValueError, match="positive definite mode"):
        PartialFactor.factorise(sym, m, FactorOptions(pivoting=PivotingMode.INDEFINITE))
