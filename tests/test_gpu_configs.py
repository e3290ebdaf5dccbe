"""The statics configurations beside the headline plate, on the GPU against the oracle and the reference's fixture.

C1 (plain FP64 statics): the 100 × 100 clamped plate of constant 2 mm thickness under 1 kPa, one load case, at its full
size (58,806 equations): D and the solution against the C oracle's multifrontal factor and solve, the centre deflection
against the clamped-plate closed form w_c = 0.00126532·q·a⁴/D_b (a sanity bound of the model, not factor parity), and
the 24 × 24 plate of the same kind against the compiled reference's fixture (tests/golden/c1plate24.npz).

C4 (batched statics): a batch of 20 plates (sides 20 to 40 elements, thickness U[1, 5] mm) factorised as one, against
the oracle's factor of the same batch; FP64 and FP32 fronts with conjugate-gradient refinement to a backward error of
1e-10, recomputed on the host with the oracle's SpMV.

Tolerances: D within 1e-9 (relative, per equation), FP64 solutions within 1e-8 of the oracle's (max norm), backward
errors on the host ≤ 1e-12 (FP64) and ≤ 1e-10 (refined FP32).
"""

import os

import numpy as np
import pytest

import _oracle
from o_o_solver_amd import (FactorOptions, FactorSolve, Precision, SupernodalFactor, SymbolicFactor, SymmetricCsc,
                            models)

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _centre_equation(w, sym):
    xy = w.coordinates.reshape(-1, 3)[:, :2]
    node = int(np.argmin(np.sum((xy - xy.mean(axis=0)) ** 2, axis=1)))
    return int(np.flatnonzero(sym.node_order == node)[0]) * w.dof_per_node + 2


def test_c1_plate_matches_the_oracle_and_the_closed_form(gpu_device):
    print("[gpu] C1 plate 100 x 100", flush=True)
    w = models.plate(100, thickness=(2e-3, 2e-3), thickness_mode="constant", pressure=1e3)
    sym = w.analyse()
    assert w.n == 58806
    b = w.loads("pressure", nrhs=1)
    panels, diag, res = _oracle.multifrontal(sym, w.matrix)
    assert res.status == 0
    xo = _oracle.oracle_solve(sym, panels, b)
    factor = SupernodalFactor.factorise(sym, w.matrix)
    assert factor.inertia.positive == w.n and factor.inertia.negative == 0 and factor.inertia.zero == 0
    assert np.max(np.abs(factor.diagonal() - diag) / np.abs(diag)) <= 1e-9
    x = FactorSolve(factor).solve(b)
    assert np.max(np.abs(x - xo)) <= 1e-8 * np.max(np.abs(xo))
    assert _oracle.backward_error(w.matrix, x, b) <= 1e-12
    # Clamped square plate under uniform pressure: w_c = 0.00126532·q·a⁴/D_b, D_b = E·t³/(12(1 − ν²)).
    bending = 70e9 * 2e-3 ** 3 / (12 * (1 - 0.3 ** 2))
    closed = 0.00126532 * 1e3 * 1.0 ** 4 / bending
    centre = float(x[_centre_equation(w, sym), 0])
    print(f"[gpu] C1 centre deflection {centre:.6e} m, closed form {closed:.6e} m", flush=True)
    assert abs(abs(centre) - closed) <= 1e-3 * closed


def test_c1_fixture_matches_the_reference(gpu_device):
    g = np.load(os.path.join(GOLDEN, "c1plate24.npz"))
    sym = SymbolicFactor(int(g["n"]), g["first_column"], g["column_count"], g["parent"], g["row_start"], g["rows"],
                         node_order=g["node_order"])
    m = SymmetricCsc(int(g["n"]), g["offsets"], g["csc_rows"], g["values"])
    factor = SupernodalFactor.factorise(sym, m)
    assert [factor.inertia.positive, factor.inertia.negative, factor.inertia.zero] == g["inertia"].tolist()
    assert np.max(np.abs(factor.diagonal() - g["diag"]) / np.abs(g["diag"])) <= 1e-9
    x = FactorSolve(factor).solve(g["b"])
    assert np.max(np.abs(x - g["x"])) <= 1e-10 * np.max(np.abs(g["x"]))


@pytest.fixture(scope="module")
def c4_batch():
    bt = models.plate_batch_c4(count=20, seed=20261205, sides=(20, 40))
    sym = bt.analyse()
    panels, diag, res = _oracle.multifrontal(sym, bt.matrix)
    assert res.status == 0
    b = bt.loads("random", nrhs=1, seed=7)
    xo = _oracle.oracle_solve(sym, panels, b)
    return bt, sym, diag, b, xo


def test_c4_batch_fp64_matches_the_oracle(gpu_device, c4_batch):
    print("[gpu] C4 batch of 20 plates, FP64", flush=True)
    bt, sym, diag, b, xo = c4_batch
    factor = SupernodalFactor.factorise(sym, bt.matrix)
    assert factor.inertia.positive == bt.n
    assert np.max(np.abs(factor.diagonal() - diag) / np.abs(diag)) <= 1e-9
    x = FactorSolve(factor).solve(b)
    assert np.max(np.abs(x - xo)) <= 1e-8 * np.max(np.abs(xo))
    assert _oracle.backward_error(bt.matrix, x, b) <= 1e-12


def test_c4_batch_fp32_fronts_refined(gpu_device, c4_batch):
    import torch
    print("[gpu] C4 batch of 20 plates, FP32 fronts + CG", flush=True)
    bt, sym, diag, b, xo = c4_batch
    f32 = SupernodalFactor.factorise(sym, bt.matrix, FactorOptions(precision=Precision.FP32))
    assert f32.inertia.positive == bt.n
    bd = torch.tensor(b, device=gpu_device)
    x, report = FactorSolve(f32).solve_refined(bd, target=1e-10, max_steps=40)
    x = x.cpu().numpy()
    eta = _oracle.backward_error(bt.matrix, x, b)
    print(f"[gpu] C4 FP32: {report}, host backward error {eta:.2e}, error vs oracle "
          f"{np.max(np.abs(x - xo)) / np.max(np.abs(xo)):.2e}", flush=True)
    assert report.converged and eta <= 1e-10
    # The forward error is bounded by about κ·η: refined further, the solution meets the oracle's.
    x, report = FactorSolve(f32).solve_refined(bd, target=1e-13, max_steps=80)
    x = x.cpu().numpy()
    assert _oracle.backward_error(bt.matrix, x, b) <= 1e-13
    assert np.max(np.abs(x - xo)) <= 1e-6 * np.max(np.abs(xo))
