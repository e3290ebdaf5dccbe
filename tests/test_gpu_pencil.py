"""The indefinite factor on the GPU (Bunch-Kaufman pivots) and the Sturm counts of shifted pencils, pinned by the
reference's inertia (golden fixture), by closed-form eigenvalues and by dense generalized eigenvalues."""

import numpy as np
import pytest
import scipy.linalg

from _problems import chain_graph, nested_chain_order
from o_o_solver_amd import (FactorOptions, FactorSolve, PivotingMode, SupernodalFactor, SymbolicFactor, SymmetricCsc,
                            models)
from o_o_solver_amd.pencil import ShiftedPencil

pytestmark = pytest.mark.gpu


def test_golden_grid_indefinite_inertia_matches_bunch_kaufman(gpu_device):
    # The reference factorised this matrix (diagonal 3.3, −1 couplings) with Bunch-Kaufman pivots; the GPU factor
    # finds its inertia, and the solutions agree.
    g = np.load(__file__.rsplit("/", 1)[0] + "/golden/grid20.npz")
    sym = SymbolicFactor(400, g["first_column"], g["column_count"], g["parent"], g["row_start"], g["rows"],
                         node_order=g["node_order"])
    m = SymmetricCsc(400, g["indefinite_offsets"], g["indefinite_csc_rows"], g["indefinite_values"])
    factor = SupernodalFactor.factorise(sym, m, FactorOptions(pivoting=PivotingMode.INDEFINITE))
    inertia = g["indefinite_inertia"].tolist()
    assert [factor.inertia.positive, factor.inertia.negative, factor.inertia.zero] == inertia == [380, 20, 0]
    x = FactorSolve(factor).solve(g["b"])
    assert np.max(np.abs(x - g["indefinite_x"])) <= 1e-10 * np.max(np.abs(g["indefinite_x"]))


def test_chain_laplacian_sturm_counts_are_closed_form(gpu_device):
    # K = tridiag(−1, 2, −1), M = I: eigenvalues 2 − 2cos(kπ/(n + 1)).
    n = 300
    g = chain_graph(n)
    sym = SymbolicFactor.analyse(g[0], g[1], g[2], nested_chain_order(n))
    position = np.empty(n, np.int64)
    position[sym.node_order] = np.arange(n)
    offsets, rows, kv, mv = [0], [], [], []
    for j in range(n):
        node = sym.node_order[j]
        entries = [(j, 2.0, 1.0)]
        for nb in (node - 1, node + 1):
            if 0 <= nb < n and position[nb] > j:
                entries.append((int(position[nb]), -1.0, 0.0))
        entries.sort()
        for r, a, b in entries:
            rows.append(r)
            kv.append(a)
            mv.append(b)
        offsets.append(len(rows))
    offsets = np.array(offsets, np.int64)
    rows = np.array(rows, np.int32)
    k = SymmetricCsc(n, offsets, rows, np.array(kv))
    mass = SymmetricCsc(n, offsets, rows, np.array(mv))
    pencil = ShiftedPencil(sym, k, mass)
    eigenvalues = 2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    for q in (3, 40, 101, 250, 299):
        shift = 0.5 * (eigenvalues[q - 1] + eigenvalues[q])
        f = pencil.factorise(shift)
        assert f.count_below == q, (shift, f.count_below)
    # The eigenvalues between the lowest and the highest Sturm point are all accounted for.
    v = pencil.sturm.verify(eigenvalues)
    assert not v.has_mismatch and v.verified_count == 299 - 3
    # σ = 2 (the midpoint of λ150 and λ151) makes every diagonal term zero: no 1 × 1 pivot in the static order would
    # do, 2 × 2 blocks and interchanges find the count.
    f = pencil.factorise(2.0)
    assert f.count_below == 150
    assert f.factor.two_by_two_pivot_count > 0


def test_plate_pencil_counts_match_dense_eigenvalues(gpu_device):
    w = models.plate(10, jitter=0.2, seed=7, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
    sym = w.analyse()
    k = w.matrix
    # A lumped mass on the stiffness's pattern: 1 on the translations, 1e-4 on the rotations, zero off the diagonal.
    columns = np.repeat(np.arange(k.n), np.diff(k.offsets))
    is_diag = k.rows == columns
    dof = np.empty(k.n, np.int64)
    for position, node in enumerate(sym.node_order):
        dof[6 * position:6 * position + 6] = np.arange(6)
    mv = np.where(is_diag, np.where(dof[columns] < 3, 1.0, 1e-4), 0.0)
    mass = SymmetricCsc(k.n, k.offsets, k.rows, mv)
    lam = scipy.linalg.eigh(k.to_dense(), mass.to_dense(), eigvals_only=True)
    pencil = ShiftedPencil(sym, k, mass)
    for q in (1, 6, 25, 80):
        shift = 0.5 * (lam[q - 1] + lam[q])
        assert (lam[q] - lam[q - 1]) > 1e-6 * lam[q]
        f = pencil.factorise(shift)
        assert f.count_below == q, (q, f.count_below)
        b = np.random.default_rng(q).standard_normal((k.n, 2))
        x = pencil.solve(f, b)
        residual = b - (k.multiply(x) - shift * mass.multiply(x))
        assert np.max(np.abs(residual)) <= 1e-8 * np.max(np.abs(b))
