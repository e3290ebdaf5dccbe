"""The device-resident ShiftedPencil in the loop it serves: a few steps of a shift-invert block Lanczos in buckling
mode (Grimes, Lewis & Simon, SIAM J. Matrix Anal. Appl. 15 (1994) §2; the reference's BlockLanczosBuckling,
src/linalg/eigen/BlockLanczosBuckling.cpp:148-205) on a stiffened cylinder, every block operation on the device:
the factor of S·(K − σB)·S formed there and kept resident, (K − σB)⁻¹ applied to blocks, the inner product of K.
The Ritz values near the shift are checked against the dense generalized eigenvalues (scipy), and the pencil's
device combination against the host's, bit for bit."""

import numpy as np
import pytest
import scipy.linalg

from o_o_solver_amd import models
from o_o_solver_amd.buckling import BucklingShifts
from o_o_solver_amd.pencil import combine, scale_symmetric

pytestmark = pytest.mark.gpu


def test_device_pencil_serves_a_block_lanczos_loop(gpu_device):
    import torch
    print("[gpu] block Lanczos on the device pencil", flush=True)
    w = models.cylinder(24, 12, stringer_every=6)
    w.analyse()
    b = BucklingShifts(w)
    pencil = b.pencil
    assert pencil.transform() == "buckling"
    k_dense, g_dense = b.stiffness.to_dense(), b.load.to_dense()
    mu = scipy.linalg.eigh(g_dense, k_dense, eigvals_only=True)
    lam = np.sort(1.0 / mu[mu > 1e-12 * np.max(mu)])
    sigma = 0.97 * lam[0]  # below the first factor, near it: the wanted values lead the transformed spectrum
    f = pencil.factorise(sigma)
    assert f.count_below == 0 and not f.factor.is_singular
    # The shifted matrix formed on the device has the host's bits: its factor equals, bit for bit, the factor of the
    # matrix formed on the host (scale_symmetric(combine(...)), itself the reference's bits, tests/test_pencil.py).
    host = scale_symmetric(combine(1.0, b.stiffness, -sigma, b.load), pencil.scale)
    from o_o_solver_amd import FactorOptions, PivotingMode, SupernodalFactor
    own = SupernodalFactor.factorise(w.symbolic, host, FactorOptions(pivoting=PivotingMode.INDEFINITE))
    assert np.array_equal(own.diagonal(), f.factor.diagonal())
    assert own.inertia == f.inertia
    del own

    def k_dot(u, v):  # uᵀ·K·v on the device
        return u.T @ pencil.multiply_inner_product(v)

    def k_orthonormal(block):
        g = k_dot(block, block)
        r = torch.linalg.cholesky(0.5 * (g + g.T))
        return torch.linalg.solve_triangular(r, block.T, upper=False).T  # block·R⁻ᵀ: Kᵀ-orthonormal columns

    p, steps = 6, 10
    q = k_orthonormal(torch.tensor(np.random.default_rng(1).standard_normal((w.n, p)), device=gpu_device))
    basis, alphas, betas = [q], [], []
    previous, beta = None, None
    for j in range(steps):
        # Op = (K − σB)⁻¹·K, self-adjoint in the inner product of K.
        wblk = pencil.solve(f, pencil.multiply_inner_product(basis[-1]))
        if previous is not None:
            wblk = wblk - previous @ beta.T
        a = k_dot(basis[-1], wblk)
        wblk = wblk - basis[-1] @ a
        for v in basis:  # full reorthogonalisation in the K inner product
            wblk = wblk - v @ k_dot(v, wblk)
        g = k_dot(wblk, wblk)
        r = torch.linalg.cholesky(0.5 * (g + g.T))
        nxt = torch.linalg.solve_triangular(r, wblk.T, upper=False).T
        alphas.append(a)
        betas.append(r.T)
        previous, beta = basis[-1], r.T
        basis.append(nxt)
    m = p * steps
    t = torch.zeros((m, m), dtype=torch.float64, device=gpu_device)
    for j in range(steps):
        t[j * p:(j + 1) * p, j * p:(j + 1) * p] = alphas[j]
        if j + 1 < steps:
            t[(j + 1) * p:(j + 2) * p, j * p:(j + 1) * p] = betas[j]
            t[j * p:(j + 1) * p, (j + 1) * p:(j + 2) * p] = betas[j].T
    nu = np.linalg.eigvalsh(0.5 * (t + t.T).cpu().numpy())
    # ν = λ / (λ − σ): the eigenvalues near the shift are the largest |ν|.
    nu = nu[np.abs(nu - 1.0) > 1e-12]
    ritz = np.sort(sigma * nu / (nu - 1.0))
    ritz = ritz[ritz > 0]
    print(f"[gpu] Ritz values {ritz[:4]}, dense {lam[:4]}", flush=True)
    assert np.allclose(ritz[:4], lam[:4], rtol=1e-9)
