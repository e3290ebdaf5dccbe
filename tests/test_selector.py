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


@pytest.mark.parametrize("soft", [False, True])
def test_restatement_gradient_matches_finite_differences(soft):
    x, y, theta = _data(1, 64)
    if soft:  # graded labels in [0, 1], as the training's measured gains are
        y = np.random.default_rng(7).uniform(0.0, 1.0, y.shape)
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
    # Graded labels in [0, 1].
    ys = np.random.default_rng(8).uniform(0.0, 1.0, y.shape)
    loss, grad = s.gradient(x, ys)
    assert abs(loss - _mlp.loss(theta32, xs, ys)) <= 1e-5
    g_ref = _mlp.gradient(theta32, xs, ys)
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


def _weights(sign, bias):
    """Parameters whose probability is monotone in the cut's depth (feature 0): hidden unit 0 = tanh(3·sign·f0 +
    bias), passed on to the output."""
    theta = np.zeros(_mlp.P)
    theta[0] = 3.0 * sign                      # W1[0, 0]
    theta[_mlp.H * _mlp.F] = bias              # b1[0]
    o = _mlp.H * _mlp.F + _mlp.H
    theta[o] = 3.0                             # W2[0, 0]
    o += _mlp.H * _mlp.H + _mlp.H
    theta[o] = 3.0                             # W3[0]
    return theta.astype(np.float32)


@pytest.mark.gpu
def test_cut_features_describe_the_levels(gpu_device):
    from o_o_solver_amd import SupernodalFactor, models
    from o_o_solver_amd.selector import PrecisionSelector
    w = models.plate(24, jitter=0.2, seed=5, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
    sym = w.analyse()
    plan = sym.plan()
    dm = plan.matrix(w.matrix)
    f = PrecisionSelector.cut_features(plan, dm)
    L = plan.levels
    assert f.shape == (L, 8) and np.all(np.isfinite(f))
    assert np.allclose(f[:, 0], np.arange(1, L + 1) / L)
    assert np.all(np.diff(f[:, 1]) >= 0) and abs(f[-1, 1] - 1.0) < 1e-6  # the flop share grows to the whole
    assert np.all(np.diff(f[:, 2]) >= 0) and np.all(np.diff(f[:, 3]) <= 0)  # orders below grow, above shrink
    assert np.all(np.diff(f[:, 6]) <= 0)  # the smallest dominance over more equations
    assert np.all(f[:, 7] == 0.0)  # no factorisation yet: no pivot growth
    SupernodalFactor.factorise(sym, w.matrix, device_matrix=dm)
    g = PrecisionSelector.cut_features(plan, dm)
    assert np.all(g[:, 7] > 0.0) and np.all(np.diff(g[:, 7]) >= 0)
    assert np.array_equal(g[:, :7], f[:, :7])


@pytest.mark.gpu
def test_learned_precision_follows_the_selector(gpu_device):
    from o_o_solver_amd import FactorOptions, Precision, SupernodalFactor, models
    from o_o_solver_amd.selector import PrecisionSelector
    w = models.plate(12, jitter=0.2, seed=5, thickness=(1e-3, 4e-3), thickness_mode="log-uniform")
    sym = w.analyse()
    plan = sym.plan()
    dm = plan.matrix(w.matrix)
    L = plan.levels
    assert L >= 3
    # p increasing in the depth: the top cut, FP32 throughout; decreasing and above 1/2: the lowest cut (one level
    # in FP32); below 1/2 everywhere: FP64 throughout.
    for sign, bias, expected, levels in ((1.0, 0.0, Precision.FP32, 0), (-1.0, 3.0, Precision.MIXED, 1),
                                         (-1.0, 0.0, Precision.FP64, 0)):
        selector = PrecisionSelector(parameters=_weights(sign, bias))
        cut, p = selector.choose(plan, dm)
        assert cut == {Precision.FP32: L, Precision.MIXED: 1, Precision.FP64: 0}[expected]
        f = SupernodalFactor.factorise(sym, w.matrix, FactorOptions(precision=Precision.LEARNED, selector=selector),
                                       device_matrix=dm)
        assert f.precision == expected and f.fp32_levels == levels
        assert (f.selected_probability >= 0.5) == (expected != Precision.FP64)


def test_cut_choice_restated_on_the_host():
    # The choice rule on the host restatement: the cut of the highest probability when it is at least 1/2.
    features = np.zeros((5, _mlp.F))
    features[:, 0] = np.arange(1, 6) / 5
    for sign, bias, cut in ((1.0, 0.0, 5), (-1.0, 3.0, 1), (-1.0, 0.0, 0)):
        p, _, _, _ = _mlp.forward(_weights(sign, bias).astype(np.float64), features)
        best = int(np.argmax(p))
        assert (best + 1 if p[best] >= 0.5 else 0) == cut


def test_trained_selector_gains_on_held_out_models():
    """The committed selector against FP64 and FP32 fronts throughout, on the models held out of its training: the
    measured time of its choices (o_o_solver_amd/train_selector.py) is at most either method's alone."""
    import os
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "o_o_solver_amd")
    with np.load(os.path.join(root, "selector_weights.npz"), allow_pickle=False) as wts:
        theta = wts["parameters"].astype(np.float64)
    with np.load(os.path.join(root, "selector_data.npz"), allow_pickle=False) as d:
        data = {k: d[k] for k in d.files}
    p, _, _, _ = _mlp.forward(theta, data["features"].astype(np.float64))
    # The host restatement gives the GPU's probabilities (FP32 kernels) to within their rounding.
    assert np.max(np.abs(p - data["probability"])) <= 1e-4
    held = np.flatnonzero(data["held_out"])
    assert held.size >= 5
    learned = fp64 = fp32 = 0.0
    for q in held:
        rows = data["problem"] == q
        best = int(np.argmax(p[rows]))
        learned += data["times"][rows][best] if p[rows][best] >= 0.5 else data["fp64_times"][q]
        fp64 += data["fp64_times"][q]
        fp32 += data["fp32_times"][q]
    assert learned <= fp64 and learned <= fp32, (learned, fp64, fp32)


@pytest.mark.gpu
def test_committed_selector_choices(gpu_device):
    """The committed weights through the GPU kernels, on a model of each kind the training measured: a thin 60 × 60
    plate keeps FP64 fronts (FP32 fronts took 3.9 ms against 2.6), the 44³ hex block with σ = 3 takes FP32 fronts
    throughout (41.3 ms against 63.3). The pivot growth comes from an FP64 factorisation first, as in training."""
    from o_o_solver_amd import FactorOptions, Precision, SupernodalFactor, models
    from o_o_solver_amd.selector import PrecisionSelector
    selector = PrecisionSelector.trained()
    assert selector is not None
    for w, expected in ((models.plate(60, jitter=0.2, seed=80, thickness=(2e-3, 2e-3), thickness_mode="constant"),
                         Precision.FP64),
                        (models.hex_block(44, seed=44, sigma=3.0), Precision.FP32)):
        sym = w.analyse()
        plan = sym.plan()
        dm = plan.matrix(w.matrix)
        SupernodalFactor.factorise(sym, w.matrix, device_matrix=dm)
        cut, p = selector.choose(plan, dm)
        print(f"[gpu] {w.name}: cut {cut} of {plan.levels}, p max {p.max():.3f}", flush=True)
        f = SupernodalFactor.factorise(sym, w.matrix, FactorOptions(precision=Precision.LEARNED, selector=selector),
                                       device_matrix=dm)
        assert f.precision == expected
        assert cut == (plan.levels if expected == Precision.FP32 else 0)
