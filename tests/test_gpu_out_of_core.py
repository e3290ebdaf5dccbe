"""The out-of-core factor (OutOfCoreFactor, include/oos/linalg/supernodal/OutOfCoreFactor.hpp:22-95) on the GPU:
panels spooled to pinned host memory level by level, streamed back by the solves.

Contract: the factor (inertia, D, every panel) and its solutions bit-equal to the in-core plan's, and so within the
in-core contract of the oracle (panels 1e-9); the spool holds every panel and inverse; the device holds two levels'
windows, less than the factor's panels."""

import numpy as np
import pytest

import _oracle
from o_o_solver_amd import (FactorOptions, FactorSolve, OutOfCoreFactor, PivotingMode, Precision, SupernodalFactor,
                            models)

pytestmark = pytest.mark.gpu

CASES = {
    "plate 40": lambda: models.plate(40, jitter=0.2, seed=6, thickness=(1e-3, 4e-3), thickness_mode="log-uniform"),
    "hex 9": lambda: models.hex_block(9, seed=3, sigma=1.5),
}


@pytest.mark.parametrize("case", list(CASES))
def test_out_of_core_factor_and_solve_equal_the_in_core_ones(gpu_device, case):
    print(f"[gpu] out of core {case}", flush=True)
    w = CASES[case]()
    sym = w.analyse()
    m = w.matrix
    ooc = OutOfCoreFactor.factorise(sym, m)
    inc = SupernodalFactor.factorise(sym, m)
    f = ooc.factor
    assert (f.inertia.positive, f.inertia.negative, f.inertia.zero) == \
        (inc.inertia.positive, inc.inertia.negative, inc.inertia.zero)
    assert np.array_equal(f.diagonal(), inc.diagonal())
    for s in range(sym.supernode_count):
        assert np.array_equal(f.panel(s)[0], inc.panel(s)[0]), s
    b = w.loads("random", nrhs=20, seed=4)
    x = ooc.solve(b)
    assert np.array_equal(x, FactorSolve(inc).solve(b))
    panels, diag, res = _oracle.multifrontal(sym, m)
    xo = _oracle.oracle_solve(sym, panels, b)
    assert np.max(np.abs(x - xo)) <= 1e-8 * np.max(np.abs(xo))
    t = ooc.traffic()
    # Every level's panels and inverses spooled once; a solve of two tiles of 16 read each level back, but for the
    # two top levels of each backward sweep.
    assert t.spooled_bytes == t.written_bytes > 8 * sym.panel_value_count - 1
    assert t.spooled_bytes <= t.read_bytes <= 4 * t.spooled_bytes
    # The device holds two windows of levels, not every panel.
    plan = sym.plan(out_of_core=True)
    assert plan.stats()["panel_values"] < sym.panel_value_count


def test_out_of_core_fp32_refined_solve_equals_the_in_core_one(gpu_device):
    import torch
    w = CASES["plate 40"]()
    sym = w.analyse()
    m = w.matrix
    options = FactorOptions(precision=Precision.FP32)
    ooc = OutOfCoreFactor.factorise(sym, m, options)
    plan = sym.plan(out_of_core=True)
    inc = SupernodalFactor.factorise(sym, m, options)
    b = torch.tensor(w.loads("random", nrhs=3, seed=2), device="cuda:0")
    x1, r1 = FactorSolve(ooc.factor).solve_refined(b, target=1e-10)
    x2, r2 = FactorSolve(inc).solve_refined(b, target=1e-10)
    assert r1.steps == r2.steps and r1.backward_error <= 1e-10
    assert torch.equal(x1, x2)
    assert _oracle.backward_error(m, x1.cpu().numpy(), b.cpu().numpy()) <= 1e-10


def test_out_of_core_refusals(gpu_device):
    w = models.plate(8, seed=1)
    sym = w.analyse()
    with pytest.raises(ValueError, match="positive definite"):
        OutOfCoreFactor.factorise(sym, w.matrix, FactorOptions(pivoting=PivotingMode.INDEFINITE))
    plan = sym.plan(out_of_core=True)
    with pytest.raises(ValueError, match="nothing to tune"):
        plan.autotune(plan.matrix(w.matrix))
