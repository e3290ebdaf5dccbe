"""The host side of shifted pencils: combine, symmetric scaling, the pencil's scales and the Sturm check
(src/linalg/sparse/SymmetricCsc.cpp:209-252, src/linalg/eigen/PencilScaling.cpp, src/linalg/eigen/SturmCheck.cpp)."""

import numpy as np
import pytest

from _problems import grid_graph, graph_matrix
from o_o_solver_amd import SymbolicFactor, SymmetricCsc
from o_o_solver_amd.pencil import (SolverError, SturmCheck, SturmPoint, combine, pencil_scale_of,
                                   scale_symmetric)


def _grid(diagonal):
    g = grid_graph(6)
    sym = SymbolicFactor.analyse(g[0], g[1], g[2], np.arange(36, dtype=np.int32))
    return sym, graph_matrix(sym, g[0], g[1], np.full(36, diagonal), lambda a, b: -1.0)


def test_combine_rounds_two_products_and_a_sum():
    _, a = _grid(4.0)
    b = SymmetricCsc(a.n, a.offsets, a.rows, np.linspace(0.1, 0.9, a.values.size))
    c = combine(0.3, a, -1.7, b)
    for k in range(a.values.size):
        left = 0.3 * float(a.values[k])
        right = -1.7 * float(b.values[k])
        assert c.values[k] == left + right
    other = SymmetricCsc(a.n, a.offsets, a.rows[::-1].copy(), a.values)
    with pytest.raises(ValueError, match="same structure"):
        combine(1.0, a, 1.0, other)


def test_pencil_scales_are_powers_of_two_that_normalise_the_diagonal():
    _, a = _grid(4.0)
    values = a.values.copy()
    diag_at = a.offsets[:-1]
    rng = np.random.default_rng(3)
    values[diag_at] = np.exp(rng.uniform(-30, 30, a.n))
    values[diag_at[5]] = -2.0      # not positive: scale 1
    values[diag_at[7]] = np.inf    # not finite: scale 1
    k = SymmetricCsc(a.n, a.offsets, a.rows, values)
    s = pencil_scale_of(k)
    mantissa, _ = np.frexp(s)
    assert np.all(mantissa == 0.5)
    assert s[5] == 1.0 and s[7] == 1.0
    scaled = scale_symmetric(k, s)
    d = scaled.values[diag_at]
    ok = np.isfinite(values[diag_at]) & (values[diag_at] > 0)
    assert np.all((d[ok] >= 0.25) & (d[ok] < 2.0))
    # Scaling by powers of two rounds nothing: scaling back gives the matrix bit for bit.
    back = scale_symmetric(scaled, 1.0 / s)
    assert np.array_equal(back.values[ok[np.repeat(np.arange(a.n), np.diff(a.offsets))]],
                          values[ok[np.repeat(np.arange(a.n), np.diff(a.offsets))]])


def test_host_multiply_is_the_dense_product():
    _, a = _grid(3.0)
    x = np.random.default_rng(1).standard_normal((a.n, 3))
    assert np.allclose(a.multiply(x), a.to_dense() @ x, rtol=1e-14, atol=1e-14)


def test_sturm_check_orders_verifies_and_refuses_contradictions():
    check = SturmCheck()
    check.add_point(SturmPoint(2.0, 5))
    check.add_point(SturmPoint(0.5, 1))
    check.add_point(SturmPoint(1.0, 3))
    check.add_point(SturmPoint(1.0, 3))  # the same point again
    assert [p.shift for p in check.points] == [0.5, 1.0, 2.0]
    assert check.count_below(1.0) == 3
    with pytest.raises(ValueError):
        check.count_below(1.5)
    with pytest.raises(SolverError):
        check.add_point(SturmPoint(1.5, 2))  # fewer below a larger shift
    with pytest.raises(SolverError):
        check.add_point(SturmPoint(1.0, 4))  # two counts at one shift
    v = check.verify([0.6, 0.9, 1.2, 1.5])
    assert not v.has_mismatch and v.verified_up_to == 2.0 and v.verified_count == 4
    v = check.verify([0.6, 1.2, 1.5])
    assert v.has_mismatch and v.verified_up_to == 0.5 and v.mismatch == (0.5, 1.0, 2, 1)


def _cylinder_pencil():
    from o_o_solver_amd import models
    w = models.cylinder(48, 20)
    w.analyse()
    return w.matrix, w.geometric()


@pytest.mark.parametrize("alpha,beta", [(1.0, -3.7), (0.3, 1e-9), (-1.0, 2.5e6)])
def test_pencil_work_equals_the_compiled_reference_bit_for_bit(reference_lib, alpha, beta):
    # combine, pencilScaleOf and scaleSymmetric against the reference's own (oracle/_ref), on a stiffened cylinder's
    # stiffness and geometric stiffness and on a grid with extreme diagonals (zero, negative, subnormal, infinite,
    # missing): every value's bits.
    import _oracle
    k, kg = _cylinder_pencil()
    combined = combine(alpha, k, beta, kg)
    assert np.array_equal(combined.values.view(np.uint64), _oracle.ref_combine(alpha, k, beta, kg).view(np.uint64))
    s = pencil_scale_of(k)
    assert np.array_equal(s.view(np.uint64), _oracle.ref_pencil_scale_of(k).view(np.uint64))
    scaled = scale_symmetric(combined, s)
    assert np.array_equal(scaled.values.view(np.uint64),
                          _oracle.ref_scale_symmetric(combined, s).view(np.uint64))
    _, g = _grid(4.0)
    values = g.values.copy()
    diag_at = g.offsets[:-1]
    values[diag_at[:6]] = [0.0, -1.0, 5e-320, np.inf, 3.0 * 2.0 ** 700, 2.0 ** -1000]
    odd = SymmetricCsc(g.n, g.offsets, g.rows, values)
    assert np.array_equal(pencil_scale_of(odd), _oracle.ref_pencil_scale_of(odd))
    # A column without its diagonal term (column 1 holds row 2 only): the scale is 1.
    missing = SymmetricCsc(3, np.array([0, 2, 3, 4]), np.array([0, 1, 2, 2], np.int32), np.array([4.0, 1.0, 1.0, 9.0]))
    assert np.array_equal(pencil_scale_of(missing), _oracle.ref_pencil_scale_of(missing))
    assert pencil_scale_of(missing)[1] == 1.0
