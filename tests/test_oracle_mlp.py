"""Pins of the warm-start network's oracle (oracle/ub_mlp_oracle.py).

Each check is fixed by something other than the oracle's own formulas:
published values, the exact-erf GELU, central finite differences, Adam's
closed-form first step, SplitMix64's published first output and torch's
bfloat16 conversion.  The product's initial parameters are checked for the
zero padding the device relies on.
"""
import math

import numpy as np
import pytest
from scipy.special import erf

import ub_mlp_oracle as M
from unblending_amd import mlp as product_mlp


def test_layout_matches_the_library():
    assert M.N_PARAMS == product_mlp.param_count() == 41376
    assert M.OFF_W == [0, 4224, 20736, 37248]
    assert M.OFF_B == [4096, 20608, 37120, 41344]


def test_gelu_known_values_and_erf_form():
    # tanh-form values: 0.5 z (1 + tanh(sqrt(2/pi)(z + 0.044715 z^3)))
    assert M.gelu(np.array(0.0)) == 0.0
    assert M.gelu(np.array(1.0)) == pytest.approx(0.8411919906082768, abs=1e-15)
    assert M.gelu(np.array(-1.0)) == pytest.approx(-0.15880800939172324, abs=1e-15)
    z = np.linspace(-6, 6, 1201)
    exact = 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))
    assert np.max(np.abs(M.gelu(z) - exact)) < 1e-3
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(M.gelu_grad(z) - fd)) < 1e-8


def _problem(seed=0, B=24, outputs=12):
    rng = np.random.default_rng(seed)
    params = product_mlp.initial_params(outputs // 4, seed).astype(np.float64)
    params += rng.normal(0, 0.05, params.size) * (params != 0)  # break the zero biases' symmetry
    W, b = M.unpack(params)
    for l in range(4):
        b[l][:] = rng.normal(0, 0.1, b[l].size)
    b[3][outputs:] = 0.0
    params = M.pack(W, b)
    images = rng.uniform(0, 1, (2, 5, 7, 3))
    idx = rng.integers(0, 2 * 35, B)
    X = M.patches(images, idx, 7, 5)
    T = rng.uniform(0, 1, (B, M.OUT_PAD))
    return params, X, T, outputs


def test_gradient_matches_central_differences():
    params, X, T, outputs = _problem()
    L, g = M.loss_grad(params, X, T, outputs)
    assert L == pytest.approx(M.loss(params, X, T, outputs), rel=1e-14)
    rng = np.random.default_rng(1)
    # indices in every weight matrix and bias, padding excluded
    picks = []
    for l, (i, o) in enumerate(M.SHAPES):
        rows = M.INPUTS if l == 0 else i
        cols = outputs if l == 3 else o
        for _ in range(12):
            r, c = rng.integers(0, rows), rng.integers(0, cols)
            picks.append(M.OFF_W[l] + r * o + c)
        for _ in range(4):
            picks.append(M.OFF_B[l] + rng.integers(0, cols))
    h = 1e-6
    for k in picks:
        p1, p2 = params.copy(), params.copy()
        p1[k] += h
        p2[k] -= h
        fd = (M.loss(p1, X, T, outputs) - M.loss(p2, X, T, outputs)) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-11), k


def test_padding_gets_zero_gradient():
    params, X, T, outputs = _problem(seed=3)
    _, g = M.loss_grad(params, X, T, outputs)
    W, b = M.unpack(g)
    assert np.all(W[0][M.INPUTS:, :] == 0.0)
    assert np.all(W[3][:, outputs:] == 0.0)
    assert np.all(b[3][outputs:] == 0.0)


def test_initial_params_padding_is_zero():
    p = product_mlp.initial_params(5, seed=7)
    W, b = M.unpack(p)
    assert np.all(W[0][M.INPUTS:, :] == 0.0) and np.all(W[3][:, 20:] == 0.0)
    assert all(np.all(x == 0.0) for x in b)
    assert np.std(W[1]) > 0.05


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(0)
    g = rng.normal(0, 1e-2, 100)
    w = rng.normal(0, 1, 100)
    w1, m, v = M.adam(w, np.zeros(100), np.zeros(100), g, 0, 1e-3)
    np.testing.assert_allclose(w1, w - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    # second step with the same gradient: the same move again
    w2, _, _ = M.adam(w1, m, v, g, 1, 1e-3)
    np.testing.assert_allclose(w2 - w1, w1 - w, rtol=1e-9, atol=1e-15)


def test_splitmix64_published_output_and_sampler():
    assert M.splitmix64(0) == 0xE220A8397B1DCDAF
    a = M.sample_indices(42, 3, 256, 1000)
    assert a.min() >= 0 and a.max() < 1000 and len(set(a.tolist())) > 200
    assert np.array_equal(a, M.sample_indices(42, 3, 256, 1000))
    assert not np.array_equal(a, M.sample_indices(42, 4, 256, 1000))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 1000),
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, 65504.0])])
    x32 = x.astype(np.float32)
    ref = torch.from_numpy(x32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(x32), ref)


def test_patches_clamp_at_edges():
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 1, (1, 4, 6, 3))
    X = M.patches(img, np.array([0, 23, 8]), 6, 4)
    for row, p in enumerate([0, 23, 8]):
        y, x = divmod(p, 6)
        k = 0
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                yy, xx = min(max(y + dy, 0), 3), min(max(x + dx, 0), 5)
                np.testing.assert_array_equal(X[row, k:k + 3], img[0, yy, xx] - 0.5)
                k += 3
        assert np.all(X[row, 27:] == 0.0)


def test_planted_mistake_fails_the_difference_check(monkeypatch):
    """A dropped term in GELU's derivative (a plausible slip) breaks the
    finite-difference agreement of the first layer's weights."""
    params, X, T, outputs = _problem(seed=5)
    monkeypatch.setattr(M, "gelu_grad", lambda z: 0.5 * (1.0 + np.tanh(M.GELU_C0 * (z + M.GELU_C1 * z ** 3))))
    _, g = M.loss_grad(params, X, T, outputs)
    bad = 0
    for k in range(M.OFF_W[0], M.OFF_W[0] + 20 * 128, 97):
        p1, p2 = params.copy(), params.copy()
        p1[k] += 1e-6
        p2[k] -= 1e-6
        fd = (M.loss(p1, X, T, outputs) - M.loss(p2, X, T, outputs)) / 2e-6
        bad += abs(g[k] - fd) > 1e-5 * abs(fd) + 1e-11
    assert bad > 5


@pytest.mark.parametrize("plant", M.PLANTS)
def test_device_gradient_bar_catches_planted_slips(plant):
    """The bar the device gradient is held to (test_gpu_mlp.GRAD_REL_TOL, per
    weight matrix and bias, relative to the oracle) is passed by the device's
    own rounding but not by a plausible derivative slip: each planted
    mistake moves some layer's gradient, on the problem the GPU test uses,
    by several times the bar."""
    from test_gpu_mlp import GRAD_REL_TOL, _perturbed_params, _train_problem
    L, B = 4, 256
    images, targets = _train_problem(L)
    count, h, w, _ = images.shape
    params = _perturbed_params(L, 3)
    idx = np.random.default_rng(4).integers(0, count * h * w, B)
    X = M.patches(images, idx, w, h)
    T = M.targets_of(targets, idx, L, w, h, 4 * L)
    _, good = M.loss_grad(params, X, T, 4 * L, bf16_ops=True)
    _, bad = M.loss_grad(params, X, T, 4 * L, bf16_ops=True, plant=plant)
    worst = 0.0
    for l, (i, o) in enumerate(M.SHAPES):
        for off, size in ((M.OFF_W[l], i * o), (M.OFF_B[l], o)):
            a, b = bad[off:off + size], good[off:off + size]
            worst = max(worst, np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    print("%s: largest relative change %.3e" % (plant, worst))
    assert worst > 5 * GRAD_REL_TOL
