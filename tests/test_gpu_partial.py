"""PartialFactor on the GPU (oosg_plan_create_partial + oosg_factorise, oosg_partial_*) against the oracle's partial
factorisation, which is bit-equal to the reference's PartialFactor (tests/test_partial.py).

Contract (FP64): K* symmetric and within 1e-10 of the oracle's (relative, max norm); the interior's inertia exact, its
D and panels within 1e-9 of the oracle's; condensed loads and recovered solutions within 1e-9 of the oracle's; the
recovery from u_b = K*⁻¹·f* solves the whole system to a backward error below 1e-13."""

import ctypes

import numpy as np
import pytest

import _oracle
from _problems import partial_plate
from o_o_solver_amd import FactorOptions, FactorSolve, PartialFactor, PivotPolicy, Precision, SupernodalFactor, _native

pytestmark = pytest.mark.gpu


def _relative(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


CASES = {"plate 12": (12, 0.2, "constant"), "plate 30": (30, 0.1, "log-uniform"), "plate 48": (48, 0.2, "constant")}


@pytest.mark.parametrize("case", list(CASES))
def test_partial_factor_matches_the_oracle(gpu_device, case):
    print(f"[gpu] partial factor {case}", flush=True)
    nx, jitter, mode = CASES[case]
    w, sym, m, t, order = partial_plate(nx, jitter=jitter, seed=3, thickness_mode=mode)
    pf = PartialFactor.factorise(sym, m)
    panels, diag, schur, res = _oracle.multifrontal_partial(sym, m)
    assert res.status == 0
    interior = sym.n - sym.trailing_dof_count
    assert (pf.dimension, pf.trailing_dof_count, pf.interior_dof_count) == (sym.n, sym.trailing_dof_count, interior)
    k = pf.schur_complement()
    assert k.shape == (sym.trailing_dof_count, sym.trailing_dof_count)
    assert np.array_equal(k, k.T)
    assert _relative(k, schur) <= 1e-10
    f_int = pf.interior_factor
    assert (f_int.inertia.positive, f_int.inertia.negative, f_int.inertia.zero) == (interior, 0, 0)
    d = f_int.diagonal()[:interior]
    assert np.max(np.abs(d - diag[:interior]) / np.abs(diag[:interior])) <= 1e-9
    offset, worst = 0, 0.0
    for s in range(sym.supernode_count - 1):
        c = int(sym.column_count[s])
        rows = c + sym.row_count(s)
        ref = panels[offset:offset + c * rows].reshape(c, rows).T
        offset += c * rows
        got, _ = f_int.panel(s)
        worst = max(worst, _relative(got, ref))
    assert worst <= 1e-9, worst
    # Loads: condensed and recovered, 20 columns (two tiles of the device's 16).
    rng = np.random.default_rng(nx)
    f = np.asfortranarray(rng.standard_normal((sym.n, 20)))
    ub = np.asfortranarray(rng.standard_normal((sym.trailing_dof_count, 20)))
    condensed = pf.condense_loads(f)
    assert condensed.shape == (sym.trailing_dof_count, 20)
    assert _relative(condensed, _oracle.oracle_condense(sym, panels, f)) <= 1e-9
    x = np.array(f, order="F", copy=True)
    out = pf.recover(x, ub)
    assert out is x  # in place
    assert np.array_equal(x[interior:], ub)
    assert _relative(x, _oracle.oracle_recover(sym, panels, f, ub)) <= 1e-9
    # The condensation is exact: u_b = K*⁻¹·f*, recovered, solves K·u = f.
    u_b = np.linalg.solve(k, condensed[:, :3])
    u = pf.recover(np.array(f[:, :3], order="F"), u_b)
    assert _oracle.backward_error(m, u, f[:, :3]) < 1e-13


def test_release_and_recompute_give_the_same_bits(gpu_device):
    w, sym, m, t, order = partial_plate(20, jitter=0.2, seed=4)
    options = FactorOptions(pivot_policy=PivotPolicy(max_ratio=1e8))
    pf = PartialFactor.factorise(sym, m, options)
    f = np.asfortranarray(np.random.default_rng(1).standard_normal((sym.n, 2)))
    before = [pf.interior_factor.panel(s)[0] for s in range(sym.supernode_count - 1)]
    condensed = pf.condense_loads(f)
    assert pf.is_factorisation_of(sym, m)
    pf.release_interior_factor()
    assert not pf.is_interior_factor_kept and not pf.is_factorisation_of(sym, m)
    with pytest.raises(RuntimeError, match="was released"):
        pf.condense_loads(f)
    with pytest.raises(RuntimeError, match="was released"):
        pf.recover(f, np.zeros((sym.trailing_dof_count, 2)))
    with pytest.raises(ValueError, match="pivot policy"):
        pf.recompute_interior_factor(sym, m, FactorOptions())
    other = type(m)(m.n, m.offsets.copy(), m.rows.copy(), m.values.copy())
    other.values[0] = np.nextafter(other.values[0], np.inf)
    with pytest.raises(ValueError, match="checksum"):
        pf.recompute_interior_factor(sym, other, options)
    pf.recompute_interior_factor(sym, m, options)
    assert pf.is_factorisation_of(sym, m)
    after = [pf.interior_factor.panel(s)[0] for s in range(sym.supernode_count - 1)]
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
    assert np.array_equal(pf.condense_loads(f), condensed)


def test_partial_factor_refusals(gpu_device):
    w, sym, m, t, order = partial_plate(12, seed=2)
    pf = PartialFactor.factorise(sym, m)
    with pytest.raises(ValueError, match="rows, expected"):
        pf.condense_loads(np.zeros((sym.n - 1, 1)))
    with pytest.raises(ValueError, match="columns of trailing values"):
        pf.recover(np.zeros((sym.n, 2)), np.zeros((sym.trailing_dof_count, 1)))
    # The whole-matrix solves and FP32 are refused on a partial plan.
    with pytest.raises(ValueError, match="partial"):
        FactorSolve(pf.interior_factor).solve(np.zeros(sym.n))
    plan = sym.plan(partial=True)
    lib = _native.lib()
    handle = ctypes.c_void_p()
    opts = SupernodalFactor._options_c(FactorOptions(), Precision.FP32, None)
    assert lib.oosg_factorise(plan.handle, plan.matrix(m).handle, ctypes.byref(opts), ctypes.byref(handle)) == 1
    assert "FP64" in lib.oosg_last_error().decode()


def test_without_a_trailing_set_the_partial_factor_is_the_whole_one(gpu_device):
    from o_o_solver_amd import models
    w = models.plate(10, jitter=0.1, seed=1)
    sym = w.analyse()
    pf = PartialFactor.factorise(sym, w.matrix)
    assert pf.trailing_dof_count == 0 and pf.schur_complement().shape == (0, 0)
    b = w.loads("random", nrhs=2, seed=3)
    assert pf.condense_loads(b).shape == (0, 2)
    x = pf.recover(np.array(b, order="F"), np.zeros((0, 2)))
    assert _oracle.backward_error(w.matrix, x, b) < 1e-13
