"""oosg_solve_refined_device_noalloc against oosg_solve_refined_device, through the C ABI: the same refinement in the
library's scratch array instead of the plan's vectors, so the same bits and reports. Two calls in a row, the second
on a smaller problem and fewer right-hand sides (the array then holds the first call's larger values), in both
precisions; a problem beyond the array is refused with OOSG_INVALID_ARGUMENT."""

import ctypes

import numpy as np
import pytest

import _oracle
import o_o_solver_amd as oos
from o_o_solver_amd import FactorOptions, Precision, SupernodalFactor, models, _native

pytestmark = pytest.mark.gpu


def _refined(entry, factor, b):
    import torch
    lib = _native.lib()
    bc = b.T.contiguous()
    xc = torch.full_like(bc, float("nan"))
    report = _native.SolveReportC()
    lib.oosg_plan_set_stream(factor.plan.handle, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    status = getattr(lib, entry)(factor.handle, factor.device_matrix.handle, bc.data_ptr(), xc.data_ptr(),
                                 bc.shape[1], bc.shape[0], 1e-10, 40, ctypes.byref(report))
    torch.cuda.synchronize()
    return status, xc.T, (report.steps, report.converged, report.backward_error)


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
def test_no_allocation_solve_equals_the_refined_solve(gpu_device, precision):
    import torch
    print(f"[gpu] no-allocation refined solve, {precision.name}", flush=True)
    for side, nrhs, seed in ((40, 20, 1), (12, 3, 2)):
        w = models.plate(side, jitter=0.2, seed=seed, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
        sym = w.analyse()
        factor = SupernodalFactor.factorise(sym, w.matrix, FactorOptions(precision=precision))
        b = w.loads("random", nrhs=nrhs, seed=seed)
        bd = torch.tensor(b, device=gpu_device)
        s0, x0, r0 = _refined("oosg_solve_refined_device", factor, bd)
        s1, x1, r1 = _refined("oosg_solve_refined_device_noalloc", factor, bd)
        assert s0 == 0 and s1 == 0, _native.lib().oosg_last_error()
        assert torch.equal(x0, x1) and r0 == r1, (side, nrhs)
        assert _oracle.backward_error(w.matrix, x1.cpu().numpy(), b) <= 1e-10
        del factor


def test_no_allocation_solve_refuses_what_does_not_fit(gpu_device):
    import torch
    w = models.plate(220)
    sym = w.analyse()
    factor = SupernodalFactor.factorise(sym, w.matrix)
    b = torch.ones((w.n, 1), dtype=torch.float64, device=gpu_device)
    status, _, _ = _refined("oosg_solve_refined_device_noalloc", factor, b)
    assert status == 1  # OOSG_INVALID_ARGUMENT
    assert "scratch" in _native.lib().oosg_last_error().decode()
