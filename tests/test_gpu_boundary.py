"""The drop-in boundary on the GPU: what the C ABI carries back besides the factor (the reference's exception payloads,
the pivot issue list, cancellation), the memory budget reserved before numeric work, the ordering of the library's
streams after the caller's, and refined solves of more than 16 right-hand sides. Every result is checked against the
oracle or the reference's known answers, never against the GPU's own figures.
"""

import numpy as np
import pytest

import _oracle
from _problems import chain_graph, graph_matrix
import o_o_solver_amd as oos
from o_o_solver_amd import FactorOptions, FactorSolve, Precision, SupernodalFactor, SymbolicFactor, models

pytestmark = pytest.mark.gpu


def _plate(side=16, seed=3):
    w = models.plate(side, jitter=0.2, seed=seed, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
    w.analyse()
    return w


def test_budget_refused_before_any_work_with_its_shortfall(gpu_device):
    w = _plate()
    sym = w.symbolic
    need = sym.estimate_memory_bytes()
    assert need > 0
    with pytest.raises(oos.MemoryBudgetError) as error:
        sym.plan(hbm_budget_bytes=need // 2)
    shortfall = error.value.shortfall
    assert shortfall.requested_bytes == need and shortfall.limit_bytes == need // 2
    assert shortfall.available_bytes == need // 2


def test_plan_reserves_what_the_estimate_says_and_a_second_factor_is_counted(gpu_device):
    w = _plate()
    sym = w.symbolic
    need = sym.estimate_memory_bytes()
    roomy = need + (256 << 20)
    plan = sym.plan(hbm_budget_bytes=roomy)
    reserved, budget = plan.memory()
    assert reserved == need and budget == roomy
    dm = plan.matrix(w.matrix)
    matrix_bytes = plan.memory()[0] - need
    assert matrix_bytes > 0
    first = SupernodalFactor.factorise(sym, w.matrix, device_matrix=dm, hbm_budget_bytes=roomy)
    assert plan.memory()[0] == need + matrix_bytes  # the factor took the storage the plan reserved
    second = SupernodalFactor.factorise(sym, w.matrix, device_matrix=dm, hbm_budget_bytes=roomy)
    storage_bytes = plan.memory()[0] - need - matrix_bytes  # a second live factor: storage of its own
    assert storage_bytes > 0
    del first, second
    # A budget that holds the plan and the matrix but not a second factor's storage refuses the second factor before
    # any numeric work; the first is unaffected, and once it is destroyed its storage serves the next one.
    tight = need + matrix_bytes + storage_bytes // 2
    tight_plan = sym.plan(hbm_budget_bytes=tight)
    dm2 = tight_plan.matrix(w.matrix)
    first = SupernodalFactor.factorise(sym, w.matrix, device_matrix=dm2, hbm_budget_bytes=tight)
    with pytest.raises(oos.MemoryBudgetError) as error:
        SupernodalFactor.factorise(sym, w.matrix, device_matrix=dm2, hbm_budget_bytes=tight)
    assert error.value.shortfall.requested_bytes == storage_bytes
    assert error.value.shortfall.limit_bytes == tight
    del first
    again = SupernodalFactor.factorise(sym, w.matrix, device_matrix=dm2, hbm_budget_bytes=tight)
    _, diag, _ = _oracle.multifrontal(sym, w.matrix)
    assert np.max(np.abs(again.diagonal() - diag) / np.abs(diag)) <= 1e-9


def _two_by_two(last_diagonal):
    g = chain_graph(2)
    sym = SymbolicFactor.analyse(g[0], g[1], g[2], np.arange(2, dtype=np.int32))
    diag = np.full(2, 2.0)
    diag[sym.node_order[1]] = last_diagonal
    return sym, graph_matrix(sym, g[0], g[1], diag, lambda a, b: 1.0)


def test_pivot_issues_and_the_dofs_of_a_singularity(gpu_device):
    # checks/linalg/FactorEdgeChecks.cpp:151-189: [[2, 1], [1, 0]] has the pivots 2 and −0.5; the second is an issue
    # (a negative pivot) with the matrix diagonal 0 and the ratio 0 / −0.5 = −0.
    sym, m = _two_by_two(0.0)
    kept = SupernodalFactor.factorise(sym, m, FactorOptions(pivot_policy=oos.PivotPolicy(bailout=-1)))
    d = kept.pivot_diagnostics
    assert d.issue_count == 1 and len(d.issues) == 1
    issue = d.issues[0]
    assert issue.equation == 1 and issue.negative_pivot
    assert issue.matrix_diagonal == 0.0 and issue.factor_diagonal == -0.5 and issue.ratio == 0.0
    with pytest.raises(oos.SingularityError) as error:
        SupernodalFactor.factorise(sym, m)
    assert error.value.dofs == [1]
    named = FactorOptions(equation_namer=lambda e: f"grid {100 + e} component 3")
    with pytest.raises(oos.SingularityError, match="grid 101 component 3") as error:
        SupernodalFactor.factorise(sym, m, named)
    assert error.value.dofs == ["grid 101 component 3"]


def test_zero_pivot_carries_its_dofs(gpu_device):
    g = chain_graph(3)
    sym = SymbolicFactor.analyse(g[0], g[1], g[2], np.arange(3, dtype=np.int32))
    m = graph_matrix(sym, g[0], g[1], np.array([2.0, 0.5, 3.0]), lambda a, b: 1.0)
    with pytest.raises(oos.SingularityError) as error:
        SupernodalFactor.factorise(sym, m)
    assert error.value.dofs == [1]


def test_fp32_pivot_failure_is_not_a_singularity_and_learned_falls_back(gpu_device):
    # [[1, b], [b, c]] with c − b² ≈ 1.2e-7 (ratio c/d ≈ 8.5e6, below MaxDiagonalRatio 1e7): positive definite in FP64,
    # but b and c rounded to FP32 give the second pivot 0 (or a negative one with a fused multiply-add).
    b, c = 1.0000000657, 1.0000002486881
    assert c - b * b > 0 and c / (c - b * b) < 1e7
    sym, _ = _two_by_two(2.0)
    m = oos.SymmetricCsc(2, np.array([0, 2, 3]), np.array([0, 1, 1], np.int32), np.array([1.0, b, c]))
    f64 = SupernodalFactor.factorise(sym, m)
    assert f64.inertia.positive == 2 and f64.pivot_diagnostics.issue_count == 0
    with pytest.raises(oos.PrecisionInsufficientError):
        SupernodalFactor.factorise(sym, m, FactorOptions(precision=Precision.FP32))
    learned = SupernodalFactor.factorise(sym, m, FactorOptions(precision=Precision.LEARNED, fp32_threshold=0.0))
    assert learned.fell_back and learned.precision == Precision.FP64
    assert np.array_equal(learned.diagonal(), f64.diagonal())


def test_cancellation_stops_a_factorisation_and_leaves_the_plan_usable(gpu_device):
    w = _plate(24, seed=4)
    sym = w.symbolic
    plan = sym.plan()
    plan.use_graphs(False)  # launched one launch at a time: polled between levels too
    polls = []

    def cancel_at_third_poll():
        polls.append(1)
        return len(polls) >= 3

    with pytest.raises(oos.OperationCancelledError):
        SupernodalFactor.factorise(sym, w.matrix, FactorOptions(cancellation=cancel_at_third_poll))
    assert len(polls) == 3
    with pytest.raises(oos.OperationCancelledError):
        SupernodalFactor.factorise(sym, w.matrix, FactorOptions(cancellation=lambda: True))
    factor = SupernodalFactor.factorise(sym, w.matrix)
    _, diag, _ = _oracle.multifrontal(sym, w.matrix)
    assert np.max(np.abs(factor.diagonal() - diag) / np.abs(diag)) <= 1e-9
    plan.use_graphs(True)


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
def test_calls_wait_for_the_callers_stream(gpu_device, precision):
    # The right-hand sides are written by torch kernels on torch's current stream just before each call (a fresh
    # tensor every time, values that change with the call): the library must read them only once they are written.
    import torch
    print(f"[gpu] caller stream, {precision.name}", flush=True)
    w = _plate(40, seed=8)
    sym = w.symbolic
    panels, diag, _ = _oracle.multifrontal(sym, w.matrix)
    base = w.loads("random", nrhs=4, seed=2)
    side = torch.cuda.Stream(device=gpu_device)
    for call in range(3):
        for stream in (torch.cuda.current_stream(gpu_device), side):
            with torch.cuda.stream(stream):
                big = torch.ones(1 << 24, device=gpu_device, dtype=torch.float64)
                for _ in range(20):  # keep the stream busy before b is written
                    big = big * 1.0000001
                bd = torch.tensor(base, device=gpu_device) * float(call + 2) + big[0] * 0.0
                factor, x, report = SupernodalFactor.factorise_and_solve(
                    sym, w.matrix, bd, FactorOptions(precision=precision), target=1e-10, max_steps=40)
                x = x.cpu().numpy()
            b = base * float(call + 2)
            xo = _oracle.oracle_solve(sym, panels, b)
            assert report.converged and _oracle.backward_error(w.matrix, x, b) <= 1e-10, (call, report)
            # A stale or half-written b would be off by a factor of order one: a loose bound on the forward error
            # (about κ·η for the FP32 factor's refinement) tells them apart.
            tolerance = 1e-8 if precision == Precision.FP64 else 1e-3
            assert np.max(np.abs(x - xo)) <= tolerance * np.max(np.abs(xo)), call
            del factor


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
def test_refined_solves_of_more_than_16_right_hand_sides(gpu_device, precision):
    import torch
    w = _plate(24, seed=9)
    sym = w.symbolic
    panels, _, _ = _oracle.multifrontal(sym, w.matrix)
    b = w.loads("random", nrhs=37, seed=3)
    factor = SupernodalFactor.factorise(sym, w.matrix, FactorOptions(precision=precision))
    x, report = FactorSolve(factor).solve_refined(torch.tensor(b, device=gpu_device), target=1e-10, max_steps=40)
    x = x.cpu().numpy()
    assert report.converged and _oracle.backward_error(w.matrix, x, b) <= 1e-10
    xo = _oracle.oracle_solve(sym, panels, b)
    tolerance = 1e-8 if precision == Precision.FP64 else 1e-3
    assert np.max(np.abs(x - xo)) <= tolerance * np.max(np.abs(xo))
    _, x2, r2 = SupernodalFactor.factorise_and_solve(sym, w.matrix, torch.tensor(b, device=gpu_device),
                                                     FactorOptions(precision=precision), target=1e-10, max_steps=40)
    assert r2.converged and _oracle.backward_error(w.matrix, x2.cpu().numpy(), b) <= 1e-10


def _liveness_bound(sym):
    """The most update-matrix values live at once when a level's update matrices are written while their children's
    are read: max over levels L of Σ r_s² over the supernodes s with level(s) <= L <= level(parent(s)) (places
    rounded to 32 values)."""
    S = sym.supernode_count
    height = np.zeros(S, np.int64)
    for s in range(S):
        if sym.parent[s] >= 0:
            height[sym.parent[s]] = max(height[sym.parent[s]], height[s] + 1)
    r = np.diff(sym.row_start)
    size = (r * r + 31) // 32 * 32
    last = np.where(sym.parent >= 0, height[np.maximum(sym.parent, 0)], height)
    levels = int(height.max()) + 1
    return max(int(size[(height <= lv) & (last >= lv)].sum()) for lv in range(levels))


@pytest.mark.parametrize("case", ["plate 60", "hex 10"])
def test_update_matrices_held_by_liveness(gpu_device, case):
    # The update matrices share one arena by liveness (a matrix's place is free once its parent's level has read
    # it): the arena is within 35 % of the most values live at once (first-fit fragmentation), below one place per
    # supernode, and the factor is the oracle's.
    w = (models.plate(60, jitter=0.2, seed=5, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
         if case.startswith("plate") else models.hex_block(10, seed=2, sigma=1.0))
    sym = w.analyse()
    arena = sym.plan().stats()["update_values"]
    bound = _liveness_bound(sym)
    every = int(np.sum(np.diff(sym.row_start) ** 2))
    print(f"[gpu] {case}: update arena {arena} values, live bound {bound}, one place each {every}", flush=True)
    assert bound <= arena <= 1.35 * bound and arena < every
    factor = SupernodalFactor.factorise(sym, w.matrix)
    _, diag, _ = _oracle.multifrontal(sym, w.matrix)
    assert np.max(np.abs(factor.diagonal() - diag) / np.abs(diag)) <= 1e-9
