"""The buckling pencil of a stiffened cylinder on the GPU: Sturm counts of K + σ·K_G (indefinite, Bunch–Kaufman)
against the compiled reference's indefinite factor on the same matrices, and against the dense generalized
eigenvalues; the roots found by bisection on the counts against the same eigenvalues."""

import numpy as np
import pytest
import scipy.linalg

import _oracle
from o_o_solver_amd import models
from o_o_solver_amd.buckling import BucklingShifts
from o_o_solver_amd.pencil import combine, scale_symmetric

pytestmark = pytest.mark.gpu


def test_cylinder_sturm_counts_match_the_reference_and_the_eigenvalues(gpu_device, reference_lib):
    w = models.cylinder(24, 12, stringer_every=6)
    sym = w.analyse()
    b = BucklingShifts(w)
    k, g = b.stiffness.to_dense(), b.load.to_dense()
    assert np.min(np.linalg.eigvalsh(g)) >= -1e-9 * np.max(np.abs(g))  # G = −K_G is positive semi-definite
    mu = scipy.linalg.eigh(g, k, eigvals_only=True)  # 1/λ: the buckling factors are the reciprocals of μ > 0
    lam = np.sort(1.0 / mu[mu > 1e-12 * np.max(mu)])
    lows, highs = b.roots(10, guess=100.0, tolerance=1e-7)
    print(f"[gpu] first buckling factors {lam[:10]}; by bisection {0.5 * (lows + highs)}", flush=True)
    assert np.all(lows <= lam[:10] * (1 + 1e-9)) and np.all(lam[:10] <= highs * (1 + 1e-9))
    roots = 0.5 * (lows + highs)
    h, _, _, _ = _oracle.ref_symbolic(w.graph_offsets, w.adjacency, w.weights, order=sym.node_order.copy())
    try:
        for sigma in BucklingShifts.shifts(roots[0], roots[-1], 8):
            f = b.pencil.factorise(sigma)
            shifted = scale_symmetric(combine(1.0, b.stiffness, -sigma, b.load), b.pencil.scale)
            rc, msg, inertia, counts, _, _, _ = _oracle.ref_factor_solve(h, shifted, pivoting=1)
            assert rc == 0, msg
            print(f"[gpu] σ = {sigma:.6g}: count {f.count_below}, reference {inertia[1]}, eigenvalues below "
                  f"{int(np.sum(lam < sigma))}; 2x2 {f.factor.two_by_two_pivot_count} ({counts[1]}), delayed "
                  f"{f.factor.delayed_pivot_count} ({counts[0]})", flush=True)
            assert f.count_below == int(inertia[1]) == int(np.sum(lam < sigma))
            assert [f.inertia.positive, f.inertia.negative, f.inertia.zero] == inertia.tolist()
            rhs = np.random.default_rng(1).standard_normal((w.n, 2))
            x = b.pencil.solve(f, rhs)
            r = rhs - (b.stiffness.multiply(x) - sigma * b.load.multiply(x))
            assert np.max(np.abs(r)) <= 1e-9 * (np.max(np.abs(k)) * np.max(np.abs(x)) + np.max(np.abs(rhs)))
    finally:
        _oracle.reference().ref_sym_free(h)
    v = b.pencil.sturm.verify(lam)
    assert not v.has_mismatch
