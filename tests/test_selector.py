"""The precision selector's MLP: the FP64 restatement pinned by central finite differences (CPU), the HIP kernels
against the restatement, training and the learned precision choice (GPU)."""

import numpy as np
import pytest

import _mlp


def _data(seed, count=256):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((count, _mlp.F))
    y = ((x[:, 0] - 0.5 * x[:, 3] + 0.25 * x[:, 6]) > 0).astype(np.float64)
    theta = rng.uniform(-0.4, 0.4, _mlp.P)
    return x, y, theta


def test_restatement_gradient_matches_finite_differences():
    x, y, theta = _data(1, 64)
    g = _mlp.gradient(theta, x, y)
    rng = np.random.default_rng(2)
    for q in rng.choice(_mlp.P, 60, replace=False).tolist() + [_mlp.P - 1, 0]:
        h = 1e-6
        e = np.zeros(_mlp.P)
        e[q] = h
        fd = (_mlp.loss(theta + e, x, y) - _mlp.loss(theta - e, x, y)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])) + 1e-9, (q, fd, g[q])


@pytest.mark.gpu
def test_selector_kernels_match_the_restatement(gpu_device):
    from o_o_solver_amd.selector import PrecisionSelector
    x, y, theta = _data(3, 300)
    s = PrecisionSelector(seed=5, parameters=theta.astype(np.float32))
    theta32 = s.parameters().astype(np.float64)
    p_ref, _, _, _ = _mlp.forward(theta32, x.astype(np.float32).astype(np.float64))
    p = s.predict(x)
    assert np.max(np.abs(p - p_ref)) <= 1e-5
    loss, grad = s.gradient(x, y)
    xs = x.astype(np.float32).astype(np.float64)
    assert abs(loss - _mlp.loss(theta32, xs, y)) <= 1e-5
    g_ref = _mlp.gradient(theta32, xs, y)
    assert np.max(np.abs(grad - g_ref)) <= 1e-5 * max(1.0, np.max(np.abs(g_ref)))


@pytest.mark.gpu
def test_selector_training_separates_a_linear_rule(gpu_device):
    from o_o_solver_amd.selector import PrecisionSelector
    x, y, _ = _data(4, 512)
    s = PrecisionSelector(seed=9)
    first, _ = s.gradient(x, y)
    last = s.train(x, y, steps=400, learning_rate=1e-2)
    assert last < 0.25 * first
    assert np.mean((s.predict(x) >= 0.5) == (y == 1)) >= 0.97
    # Deterministic: the same seed and data give the same parameters bit for bit.
    t = PrecisionSelector(seed=9)
    t.train(x, y, steps=400, learning_rate=1e-2)
    assert np.array_equal(s.parameters(), t.parameters())


@pytest.mark.gpu
def test_learned_precision_follows_the_selector(gpu_device):
    from o_o_solver_amd import FactorOptions, Precision, SupernodalFactor, models
    from o_o_solver_amd.selector import PARAMETERS, PrecisionSelector
    w = models.plate(12, jitter=0.2, seed=5, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
    sym = w.analyse()
    features = PrecisionSelector.features(sym.plan(), sym.plan().matrix(w.matrix))
    assert features.shape == (8,) and np.all(np.isfinite(features))
    for bias, expected in ((6.0, Precision.FP32), (-6.0, Precision.FP64)):
        theta = np.zeros(PARAMETERS, np.float32)
        theta[-1] = bias
        f = SupernodalFactor.factorise(sym, w.matrix, FactorOptions(precision=Precision.LEARNED,
                                                                     selector=PrecisionSelector(parameters=theta)))
        assert f.precision == expected
        assert (f.selected_probability >= 0.5) == (bias > 0)
