"""The oracle (oracle/oracle.c) pinned: bit-equal to the compiled reference's kernel family and multifrontal factor,
and checked against closed forms, reconstructions and the known answers of the reference's own checks."""

import ctypes
import os

import numpy as np
import pytest

import _oracle
from _oracle import P
from _problems import chain_graph, graph_matrix, grid_graph, random_spd
from o_o_solver_amd import SymbolicFactor, SymmetricCsc, models, nested_dissection

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _split(front, c):
    """(panel n × c, contribution r × r) column-major copies of a dense front."""
    n = front.shape[0]
    panel = np.asfortranarray(np.tril(front)[:, :c].copy())
    contrib = np.asfortranarray(np.tril(front)[c:, c:].copy())
    return panel, contrib, n - c


@pytest.mark.parametrize("c,r", [(1, 0), (5, 3), (37, 29), (100, 70), (130, 0)])
def test_front_nopivot_bits_equal_reference(reference_lib, c, r):
    a = random_spd(c + r, seed=c * 7 + r)
    p1, c1, _ = _split(a, c)
    p2, c2, _ = _split(a, c)
    e1 = _oracle.oracle().orc_front_nopivot(P(p1), c + r, P(c1) if r else None, max(r, 1), c, r)
    e2 = reference_lib.ref_dense_front_nopivot(P(p2), c + r, P(c2) if r else None, max(r, 1), c, r)
    assert e1 == e2 == c
    assert np.array_equal(p1, p2)
    assert np.array_equal(c1, c2)


def test_front_nopivot_reconstructs():
    n, c = 60, 60
    a = random_spd(n, seed=3)
    panel, _, _ = _split(a, c)
    assert _oracle.oracle().orc_front_nopivot(P(panel), n, None, 1, c, 0) == c
    L = np.tril(panel, -1) + np.eye(n)
    D = np.diag(np.diag(panel))
    assert np.linalg.norm(L @ D @ L.T - a) / np.linalg.norm(a) < 1e-14


def test_tridiagonal_closed_form():
    # [-1, 2, -1]: d_k = (k + 1)/k, L(k + 1, k) = −k/(k + 1) (1-based).
    n = 50
    a = np.asfortranarray(2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1))
    assert _oracle.oracle().orc_ldlt_nopivot(P(a), n, n, n) == n
    k = np.arange(1, n + 1, dtype=float)
    assert np.allclose(np.diag(a), (k + 1) / k, rtol=1e-14, atol=0)
    assert np.allclose(np.diag(a, -1), -k[:-1] / (k[:-1] + 1), rtol=1e-14, atol=0)


def test_zero_pivot_stops_at_its_column(reference_lib):
    # [[2, 1, 0], [1, 0.5, 1], [0, 1, 3]]: the second pivot is 0.5 − 1/2 = 0 exactly.
    a = np.asfortranarray(np.array([[2, 1, 0], [1, 0.5, 1], [0, 1, 3]], float))
    b = a.copy(order="F")
    assert _oracle.oracle().orc_ldlt_nopivot(P(a), 3, 3, 3) == 1
    panel = np.asfortranarray(np.tril(b))
    assert reference_lib.ref_dense_front_nopivot(P(panel), 3, None, 1, 3, 0) == 1


def _bk(lib_fn, a, candidates):
    a = np.asfortranarray(np.tril(a).copy())
    n = a.shape[0]
    perm = np.zeros(candidates, np.int64)
    sub = np.zeros(candidates)
    e = lib_fn(P(a), n, n, candidates, P(perm), P(sub))
    return e, a, perm, sub


@pytest.mark.parametrize("n,candidates,seed", [(40, 40, 1), (40, 25, 2), (12, 12, 3), (30, 10, 4)])
def test_bunch_kaufman_bits_equal_reference(reference_lib, n, candidates, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    a = a + a.T
    # Small diagonals force 2 × 2 pivots and delays.
    a[np.diag_indices(n)] *= 0.05
    r1 = _bk(_oracle.oracle().orc_bunch_kaufman, a, candidates)
    r2 = _bk(reference_lib.ref_dense_bunch_kaufman, a, candidates)
    assert r1[0] == r2[0]
    for x, y in zip(r1[1:], r2[1:]):
        assert np.array_equal(x, y)


def test_bunch_kaufman_reconstructs_and_counts_inertia():
    rng = np.random.default_rng(11)
    n = 30
    a = rng.standard_normal((n, n))
    a = a + a.T
    e, f, perm, sub = _bk(_oracle.oracle().orc_bunch_kaufman, a, n)
    assert e == n
    L = np.tril(f, -1) + np.eye(n)
    D = np.diag(np.diag(f)) + np.diag(sub[:-1], -1) + np.diag(sub[:-1], 1)
    pa = a[np.ix_(perm, perm)]
    assert np.linalg.norm(L @ D @ L.T - pa) / np.linalg.norm(a) < 1e-12
    inertia = np.zeros(3, np.int64)
    _oracle.oracle().orc_inertia(P(np.ascontiguousarray(np.diag(f))), P(sub), n, P(inertia))
    eig = np.linalg.eigvalsh(a)
    assert inertia.tolist() == [int((eig > 0).sum()), int((eig < 0).sum()), 0]


def test_bunch_kaufman_two_by_two_inertia():
    a = np.array([[0.0, 1.0], [1.0, 0.0]])
    e, f, perm, sub = _bk(_oracle.oracle().orc_bunch_kaufman, a, 2)
    assert e == 2 and sub[0] == 1.0
    inertia = np.zeros(3, np.int64)
    _oracle.oracle().orc_inertia(P(np.ascontiguousarray(np.diag(f))), P(sub), 2, P(inertia))
    assert inertia.tolist() == [1, 1, 0]
    d = np.array([1.0, -2.0, 3.0, -4.0, 5.0])
    e, f, perm, sub = _bk(_oracle.oracle().orc_bunch_kaufman, np.diag(d), 5)
    inertia = np.zeros(3, np.int64)
    _oracle.oracle().orc_inertia(P(np.ascontiguousarray(np.diag(f))), P(sub), 5, P(inertia))
    assert inertia.tolist() == [3, 2, 0]


@pytest.fixture(scope="module")
def plate_problem():
    w = models.plate(12, jitter=0.2, seed=5, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
    w.analyse()
    return w


def test_spmv_bits_equal_reference(reference_lib, plate_problem):
    m = plate_problem.matrix
    x = np.random.default_rng(1).standard_normal(m.n)
    y1 = _oracle.oracle_spmv(m, x)
    y2 = np.zeros(m.n)
    err = ctypes.create_string_buffer(256)
    assert reference_lib.ref_spmv(m.n, P(m.offsets), P(m.rows), P(m.values), P(x), P(y2), err, 256) == 0
    assert np.array_equal(y1, y2)
    dense = m.to_dense()
    assert np.allclose(y1, dense @ x, rtol=1e-12, atol=1e-12 * np.abs(dense).max() * np.abs(x).max())


def _ref_factor(w, sym, matrix, b=None, kernel=0):
    h, sizes, flops, arr = _oracle.ref_symbolic(w[0], w[1], w[2], order=sym.node_order.copy())
    try:
        return _oracle.ref_factor_solve(h, matrix, b=b, kernel=kernel, panel_values=int(sizes[4]))
    finally:
        _oracle.reference().ref_sym_free(h)


def test_multifrontal_bits_equal_reference_plate(reference_lib, plate_problem):
    w = plate_problem
    sym, m = w.symbolic, w.matrix
    panels, diag, res = _oracle.multifrontal(sym, m)
    assert res.status == 0
    b = w.loads("random", nrhs=2, seed=3)
    bref = b.copy(order="F")
    rc, msg, inertia, counts, times, rpanels, rdiag = _ref_factor(
        (w.graph_offsets, w.adjacency, w.weights), sym, m, b=bref)
    assert rc == 0, msg
    assert np.array_equal(panels, rpanels)
    assert np.array_equal(diag, rdiag)
    assert list(res.inertia) == inertia.tolist() == [m.n, 0, 0]
    x = _oracle.oracle_solve(sym, panels, b)
    assert np.max(np.abs(x - bref)) <= 1e-12 * np.max(np.abs(bref))


def test_multifrontal_bits_equal_reference_grid(reference_lib):
    g = grid_graph(30)
    # The reference's own minimum-degree order (checks/linalg/FactorOrderChecks.cpp:76-79 orders its grid so).
    h, sizes, flops, arr = _oracle.ref_symbolic(g[0], g[1], g[2], method=1)
    reference_lib.ref_sym_free(h)
    sym = SymbolicFactor.analyse(g[0], g[1], g[2], arr["node_order"])
    m = graph_matrix(sym, g[0], g[1], np.full(900, 4.5), lambda a, b: -1.0)
    panels, diag, res = _oracle.multifrontal(sym, m)
    rc, msg, inertia, counts, times, rpanels, rdiag = _ref_factor(g, sym, m)
    assert rc == 0 and res.status == 0
    assert np.array_equal(panels, rpanels)
    # The multifrontal factor is the dense LDLᵀ of the matrix in factor order.
    dense = np.asfortranarray(m.to_dense())
    assert _oracle.oracle().orc_ldlt_nopivot(P(dense), m.n, m.n, m.n) == m.n
    assert np.allclose(diag, np.diag(dense), rtol=1e-13, atol=0)


@pytest.mark.parametrize("last_diagonal,negatives,issues", [(0.0, 1, 1), (-1.0, 1, 1), (2.0, 0, 0)])
def test_two_by_two_known_answers(last_diagonal, negatives, issues):
    # checks/linalg/FactorEdgeChecks.cpp:151-189: [[2, 1], [1, d]]; d ∈ {0, −1} give one negative pivot and one
    # NegativePivot issue at equation 1; d = 2 gives none.
    g = chain_graph(2)
    sym = SymbolicFactor.analyse(g[0], g[1], g[2], np.arange(2, dtype=np.int32))
    diag = np.full(2, 2.0)
    diag[sym.node_order[1]] = last_diagonal
    m = graph_matrix(sym, g[0], g[1], diag, lambda a, b: 1.0)
    panels, d, res = _oracle.multifrontal(sym, m)
    assert res.status == 0
    assert res.inertia[1] == negatives
    assert res.issue_count == issues
    if issues:
        assert res.first_issue == 1 and res.first_issue_negative == 1


def test_two_by_two_known_answers_match_reference(reference_lib):
    g = chain_graph(2)
    for last in (0.0, -1.0, 2.0):
        sym = SymbolicFactor.analyse(g[0], g[1], g[2], np.arange(2, dtype=np.int32))
        diag = np.full(2, 2.0)
        diag[sym.node_order[1]] = last
        m = graph_matrix(sym, g[0], g[1], diag, lambda a, b: 1.0)
        _, _, res = _oracle.multifrontal(sym, m)
        h, sizes, flops, arr = _oracle.ref_symbolic(g[0], g[1], g[2], order=np.arange(2, dtype=np.int32))
        rc, msg, inertia, counts, *_ = _oracle.ref_factor_solve(h, m, bailout=-1)
        reference_lib.ref_sym_free(h)
        assert rc == 0
        assert list(res.inertia) == inertia.tolist()
        assert res.issue_count == counts[2]
        assert res.first_issue == counts[3]


def test_zero_pivot_reports_singular_equation():
    # A chain whose second equation has a zero pivot: status 3 at equation 1 (SingularityError).
    g = chain_graph(3)
    sym = SymbolicFactor.analyse(g[0], g[1], g[2], np.arange(3, dtype=np.int32))
    m = graph_matrix(sym, g[0], g[1], np.array([2.0, 0.5, 3.0]), lambda a, b: 1.0)
    _, _, res = _oracle.multifrontal(sym, m)
    assert res.status == 3 and res.zero_pivot_count == 1 and res.first_zero_pivot == 1


def test_solve_manufactured_solution(plate_problem):
    w = plate_problem
    sym, m = w.symbolic, w.matrix
    panels, _, res = _oracle.multifrontal(sym, m)
    xs = np.random.default_rng(9).standard_normal((m.n, 3))
    dense = m.to_dense()
    b = dense @ xs
    x = _oracle.oracle_solve(sym, panels, b)
    cond = np.linalg.cond(dense)
    assert np.max(np.abs(x - xs)) / np.max(np.abs(xs)) < 10 * cond * 2.2e-16
    r = b - dense @ x
    eta = np.max(np.abs(r)) / (np.abs(dense).sum(1).max() * np.abs(x).max() + np.abs(b).max())
    assert eta < 1e-15


def test_oracle_solves_the_c1_fixture_and_meets_the_closed_form():
    # The clamped 24 × 24 plate of constant thickness under 1 kPa (tests/golden/c1plate24.npz, the compiled
    # reference's solution): the oracle's factor and solve give the reference's solution, and the centre deflection is
    # the clamped-plate closed form w_c = 0.00126532·q·a⁴/D_b within 0.5 % (a coarse mesh of the model).
    g = np.load(os.path.join(GOLDEN, "c1plate24.npz"))
    sym = SymbolicFactor(int(g["n"]), g["first_column"], g["column_count"], g["parent"], g["row_start"], g["rows"],
                         node_order=g["node_order"])
    m = SymmetricCsc(int(g["n"]), g["offsets"], g["csc_rows"], g["values"])
    panels, diag, res = _oracle.multifrontal(sym, m)
    assert res.status == 0 and list(res.inertia) == g["inertia"].tolist()
    assert np.array_equal(diag, g["diag"])
    x = _oracle.oracle_solve(sym, panels, g["b"])
    assert np.max(np.abs(x - g["x"])) <= 1e-12 * np.max(np.abs(g["x"]))
    closed = 0.00126532 * 1e3 / (70e9 * 2e-3 ** 3 / (12 * (1 - 0.3 ** 2)))
    assert abs(abs(float(g["x"][int(g["centre"]), 0])) - closed) <= 5e-3 * closed
