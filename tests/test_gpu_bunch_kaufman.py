"""Bunch–Kaufman pivoting on the GPU (indefinite mode) against the compiled reference's own indefinite factor
(MultifrontalDriver::eliminatePivoted, MultifrontalDriver.cpp:346-436) and the oracle's dense Bunch–Kaufman
(DenseBlockKernels.cpp:448-580): the same inertia, the same pivots (D per equation, 2 × 2 and delayed counts) where
the pivot tests are not within rounding of their thresholds, and solutions with a small backward error."""

import numpy as np
import pytest

import _oracle
from _problems import graph_matrix
from o_o_solver_amd import FactorOptions, FactorSolve, PivotingMode, SupernodalFactor, SymbolicFactor, SymmetricCsc, \
    models
from o_o_solver_amd.pencil import combine, pencil_scale_of, scale_symmetric

pytestmark = pytest.mark.gpu

INDEFINITE = FactorOptions(pivoting=PivotingMode.INDEFINITE)


def backward_error(a, x, b):
    r = b - a.multiply(x)
    norm = np.max(np.sum(np.abs(a.to_dense()), axis=1)) if a.n <= 3000 else None
    if norm is None:
        # ‖A‖∞ from the lower CSC: row sums of |a_ij| over both triangles.
        cols = np.repeat(np.arange(a.n), np.diff(a.offsets))
        sums = np.zeros(a.n)
        np.add.at(sums, a.rows, np.abs(a.values))
        np.add.at(sums, cols, np.where(a.rows != cols, np.abs(a.values), 0.0))
        norm = np.max(sums)
    return np.max(np.abs(r)) / (norm * np.max(np.abs(x)) + np.max(np.abs(b)))


def reference_indefinite(graph, order, matrix, b):
    """The compiled reference's indefinite factor and solve on the same structure: (inertia, counts, D by equation,
    x)."""
    h, _, _, _ = _oracle.ref_symbolic(graph[0], graph[1], graph[2], order=np.array(order, np.int32))
    try:
        x = np.array(b, order="F", copy=True)
        rc, msg, inertia, counts, _, _, diag = _oracle.ref_factor_solve(h, matrix, b=x, pivoting=1)
    finally:
        _oracle.reference().ref_sym_free(h)
    assert rc == 0, msg
    return inertia, counts, diag, x


def plate_pencil(nx, fraction, seed=3):
    """S·(K − σM)·S of a jittered plate with a lumped mass (1 on translations, 1e-4 on rotations), σ a fraction of the
    median diagonal ratio K_jj / M_jj."""
    w = models.plate(nx, jitter=0.2, seed=seed, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
    sym = w.analyse()
    k = w.matrix
    cols = np.repeat(np.arange(k.n), np.diff(k.offsets))
    diag = k.rows == cols
    mv = np.where(diag, np.where((k.rows % 6) < 3, 1.0, 1e-4), 0.0)
    mass = SymmetricCsc(k.n, k.offsets, k.rows, mv)
    dk, dm = k.values[diag], mv[diag]
    shift = fraction * np.median(dk / dm)
    return w, sym, scale_symmetric(combine(1.0, k, -shift, mass), pencil_scale_of(k))


@pytest.mark.parametrize("nx,fraction", [(12, 0.01), (12, 1.0), (30, 1e-3), (30, 1.0), (60, 0.5)])
def test_plate_pencil_pivots_match_the_reference(gpu_device, reference_lib, nx, fraction):
    w, sym, a = plate_pencil(nx, fraction)
    b = np.random.default_rng(nx).standard_normal((a.n, 3))
    inertia, counts, diag, x_ref = reference_indefinite((w.graph_offsets, w.adjacency, w.weights), sym.node_order, a,
                                                        b)
    f = SupernodalFactor.factorise(sym, a, INDEFINITE)
    got = [f.inertia.positive, f.inertia.negative, f.inertia.zero]
    print(f"plate {nx} σ-fraction {fraction}: inertia {got} (reference {inertia.tolist()}), delayed "
          f"{f.delayed_pivot_count} ({counts[0]}), 2x2 {f.two_by_two_pivot_count} ({counts[1]})", flush=True)
    assert got == inertia.tolist()
    # The same pivot sequence as the reference's (the same delayed columns and 2 × 2 pivots), so D per equation
    # agrees to rounding.
    assert (f.delayed_pivot_count, f.two_by_two_pivot_count) == (int(counts[0]), int(counts[1]))
    scale = np.maximum(np.abs(diag), 1e-300)
    assert np.max(np.abs(f.diagonal() - diag) / scale) <= 1e-8
    x = FactorSolve(f).solve(b)
    assert backward_error(a, x, b) <= 1e-12
    assert np.max(np.abs(x - x_ref)) <= 1e-6 * np.max(np.abs(x_ref))


def clique_graph(groups, links):
    """Scalar nodes in cliques (lists of nodes), plus the edges between groups listed in `links` (every node of one
    group to every node of the other)."""
    n = sum(len(g) for g in groups)
    adjacency = [set() for _ in range(n)]
    for g in groups:
        for a in g:
            adjacency[a].update(v for v in g if v != a)
    for i, j in links:
        for a in groups[i]:
            for v in groups[j]:
                adjacency[a].add(v)
                adjacency[v].add(a)
    offsets = np.cumsum([0] + [len(s) for s in adjacency]).astype(np.int64)
    nb = np.concatenate([sorted(s) for s in adjacency]).astype(np.int32)
    return offsets, nb, np.ones(n, np.int32)


@pytest.mark.parametrize("sizes,seed", [((5, 7, 6), 1), ((20, 30, 16), 2), ((70, 90, 40), 3), ((150, 130, 110), 4)])
def test_weak_children_delay_their_columns_to_the_parent(gpu_device, reference_lib, sizes, seed):
    # Two cliques A and B joined only through a separator S (order A, B, S: supernodes A and B, parent S). A's and B's
    # own couplings are tiny against their couplings to S, so their columns fail the pivot tests and are delayed
    # into S's front, where they are eliminated with S's columns.
    a_n, b_n, s_n = sizes
    groups = [list(range(a_n)), list(range(a_n, a_n + b_n)), list(range(a_n + b_n, a_n + b_n + s_n))]
    graph = clique_graph(groups, [(0, 2), (1, 2)])
    order = np.arange(a_n + b_n + s_n, dtype=np.int32)
    sym = SymbolicFactor.analyse(graph[0], graph[1], graph[2], order)
    rng = np.random.default_rng(seed)
    n = a_n + b_n + s_n
    group_of = np.repeat([0, 1, 2], sizes)
    values = {}

    def coupling(u, v):
        if (u, v) not in values:
            weak = group_of[u] == group_of[v] and group_of[u] < 2
            values[(u, v)] = rng.standard_normal() * (1e-4 if weak else 1.0)
        return values[(u, v)]

    diagonal = np.where(group_of < 2, 1e-5, 1.0) * rng.standard_normal(n)
    m = graph_matrix(sym, graph[0], graph[1], diagonal, coupling)
    b = rng.standard_normal((n, 2))
    inertia, counts, diag, x_ref = reference_indefinite(graph, order, m, b)
    f = SupernodalFactor.factorise(sym, m, INDEFINITE)
    got = [f.inertia.positive, f.inertia.negative, f.inertia.zero]
    print(f"cliques {sizes}: inertia {got} ({inertia.tolist()}), delayed {f.delayed_pivot_count} ({counts[0]}), "
          f"2x2 {f.two_by_two_pivot_count} ({counts[1]})", flush=True)
    assert got == inertia.tolist()
    assert counts[0] > 0 and f.delayed_pivot_count > 0
    assert f.delayed_pivot_count == counts[0] and f.two_by_two_pivot_count == counts[1]
    assert np.max(np.abs(f.diagonal() - diag) / np.maximum(np.abs(diag), 1e-300)) <= 1e-8
    # The parent's panel holds the delayed equations among its rows.
    root = sym.supernode_count - 1
    values_root, rows_root = f.panel(root)
    assert values_root.shape == (n - sum(f.panel(s)[0].shape[1] for s in range(root)), values_root.shape[1])
    assert len(set(rows_root.tolist())) == len(rows_root)
    x = FactorSolve(f).solve(b)
    assert backward_error(m, x, b) <= 1e-12
    assert np.max(np.abs(x - x_ref)) <= 1e-8 * np.max(np.abs(x_ref))


@pytest.mark.parametrize("n,seed", [(3, 1), (40, 2), (130, 3), (257, 4)])
def test_dense_front_matches_the_oracle_bunch_kaufman(gpu_device, n, seed):
    # A complete graph is one supernode, its front the whole matrix: the GPU's blocked, left-looking elimination
    # against the oracle's right-looking one (the reference's order), pivot for pivot.
    graph = clique_graph([list(range(n))], [])
    sym = SymbolicFactor.analyse(graph[0], graph[1], graph[2], np.arange(n, dtype=np.int32))
    assert sym.supernode_count == 1
    rng = np.random.default_rng(seed)
    dense = rng.standard_normal((n, n))
    dense = dense + dense.T
    dense[np.arange(n), np.arange(n)] *= 1e-3  # small diagonals: 2 × 2 pivots and interchanges
    assert np.array_equal(sym.node_order, np.arange(n))
    rows = np.concatenate([np.arange(j, n) for j in range(n)]).astype(np.int32)
    offsets = np.concatenate([[0], np.cumsum(np.arange(n, 0, -1))]).astype(np.int64)
    m = SymmetricCsc(n, offsets, rows, np.concatenate([dense[j:, j] for j in range(n)]))
    f = SupernodalFactor.factorise(sym, m, INDEFINITE)
    a = np.array(np.tril(dense), order="F")
    perm = np.zeros(n, np.int64)
    sub = np.zeros(n)
    e = _oracle.oracle().orc_bunch_kaufman(_oracle.P(a), n, n, n, _oracle.P(perm), _oracle.P(sub))
    assert e == n
    values, rows = f.panel(0)
    assert rows.tolist() == perm.tolist()
    assert np.allclose(f.subdiagonal(0), sub, rtol=1e-9, atol=0)
    scale = np.max(np.abs(a))
    assert np.max(np.abs(np.tril(values, -1) - np.tril(a, -1))) <= 1e-9 * max(1.0, np.max(np.abs(np.tril(a, -1))))
    assert np.max(np.abs(np.diag(values) - np.diag(a))) <= 1e-10 * scale
    b = rng.standard_normal((n, 2))
    x = FactorSolve(f).solve(b)
    assert np.max(np.abs(dense @ x - b)) <= 1e-9 * (np.max(np.abs(dense)) * np.max(np.abs(x)) + np.max(np.abs(b)))
