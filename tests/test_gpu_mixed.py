"""Mixed-precision factors (OOSG_MIXED): FP32 fronts in the lowest levels of the supernode tree (the subtrees below a
cut), FP64 above, on the diagonally scaled matrix, refined by conjugate gradients.

Contract: the FP32 levels' panels are the FP32 factor's bit for bit, and with the cut at the top the whole factor and
its refined solution are the FP32 factor's; every cut refines to a backward error of 1e-10 or less (recomputed on the
host with the oracle's product) in no more solves than the FP32 factor; mixed factors are refused in indefinite mode."""

import numpy as np
import pytest
import torch

import _oracle
from o_o_solver_amd import FactorOptions, FactorSolve, PivotingMode, Precision, SupernodalFactor, models

pytestmark = pytest.mark.gpu


def _plate():
    return models.plate(60, jitter=0.2, seed=9, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")


def test_mixed_cuts_refine_to_the_target(gpu_device):
    w = _plate()
    sym = w.analyse()
    m = w.matrix
    plan = sym.plan()
    levels = plan.stats()["levels"]
    b = torch.tensor(w.loads("random", nrhs=4, seed=3), device="cuda:0")
    f32 = SupernodalFactor.factorise(sym, m, FactorOptions(precision=Precision.FP32))
    x32, r32 = FactorSolve(f32).solve_refined(b, target=1e-10, max_steps=40)
    panels32 = [f32.panel(s)[0] for s in range(sym.supernode_count)]
    level_of = np.zeros(sym.supernode_count, np.int64)
    for s in range(sym.supernode_count):
        if sym.parent[s] >= 0:
            level_of[sym.parent[s]] = max(level_of[sym.parent[s]], level_of[s] + 1)
    del f32
    for cut in (levels, levels - 2, levels // 2, 0):
        print(f"[gpu] mixed cut {cut} of {levels}", flush=True)
        f = SupernodalFactor.factorise(sym, m, FactorOptions(precision=Precision.MIXED, fp32_levels=cut))
        assert f.precision == Precision.MIXED and f.fp32_levels == min(cut, levels)
        assert (f.inertia.positive, f.inertia.negative) == (sym.n, 0)
        for s in range(sym.supernode_count):
            if level_of[s] < cut:
                assert np.array_equal(f.panel(s)[0], panels32[s]), s
        x, r = FactorSolve(f).solve_refined(b, target=1e-10, max_steps=40)
        assert r.converged and r.backward_error <= 1e-10
        assert _oracle.backward_error(m, x.cpu().numpy(), b.cpu().numpy()) <= 1e-10
        assert r.steps <= r32.steps
        if cut >= levels:
            assert torch.equal(x, x32)
        del f


def test_mixed_refusals(gpu_device):
    w = models.plate(8, seed=1)
    sym = w.analyse()
    with pytest.raises(ValueError, match="mixed"):
        SupernodalFactor.factorise(sym, w.matrix, FactorOptions(precision=Precision.MIXED,
                                                                pivoting=PivotingMode.INDEFINITE))
