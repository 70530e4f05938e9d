"""GPU checks of the warm-start network and the warm-started first pass.

The network runs BF16 matrix-core products with FP32 accumulation; its
oracle (oracle/ub_mlp_oracle.py) rounds the same operands to BF16 and
accumulates in FP64, so the bars are:
  forward    max |y_gpu - y_oracle| <= 2e-3 on the sigmoid outputs
  gradient   per weight matrix / bias, |g_gpu - g_oracle| / |g_oracle| <= 2e-3
             (measured up to 1.8e-4: with the operands rounded alike, the
             device's FP32 accumulation and the odd BF16 rounding flip are the
             only differences; the derivative slips of
             test_oracle_mlp.py::test_device_gradient_bar_catches_planted_slips
             move some layer's gradient by 0.39 to 3.1)
  loss       relative 1e-3
  Adam       the FP32 update of the returned gradient within 1e-6 absolute
  sampling   the batch the device draws equals the oracle's SplitMix64 batch
             (checked through the gradient, which any other batch changes)
The warm-started first pass is checked against the oracle's solver started
from the same predicted points, pixel by pixel, at the first-pass bar
(1e-3 max abs; the expected difference is zero): a pixel whose warm attempt
has not converged within the warm outer cap is taken from the oracle's plain
solve, as the device solves it again from the default start.
"""
import numpy as np
import pytest

import ub_mlp_oracle as M
from conftest import layer_infos, oracle_layers

pytestmark = pytest.mark.gpu

FORWARD_TOL = 2e-3
GRAD_REL_TOL = 2e-3
LOSS_REL_TOL = 1e-3
ADAM_TOL = 1e-6
ALPHA_COLOR_TOL = 1e-3
WARM_OUTER_CAP = 8  # the library's default warm_outer_cap


def _perturbed_params(L, seed):
    from unblending_amd import mlp
    rng = np.random.default_rng(seed)
    p = mlp.initial_params(L, seed).astype(np.float64)
    W, b = M.unpack(p)
    for l in range(4):
        b[l][:] = rng.normal(0, 0.1, b[l].size)
    b[3][4 * L:] = 0.0
    return M.pack(W, b).astype(np.float32)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_forward_matches_oracle():
    from unblending_amd import mlp
    L = 5
    params = _perturbed_params(L, 1)
    model = mlp.WarmStartModel(L, params=params)
    rng = np.random.default_rng(0)
    rgb = rng.uniform(0, 1, (23, 37, 3))  # 851 pixels: the last 128-pixel tile is partial
    print("case forward 37x23 L=5", flush=True)
    y = model.predict(_dev(rgb)).cpu().numpy()
    ref = M.predict(params, rgb, L)
    d = np.abs(y - ref)
    print("forward max|diff| %.3e mean %.3e" % (d.max(), d.mean()), flush=True)
    assert d.max() <= FORWARD_TOL


def test_forward_across_prediction_chunks():
    """A large image (280000 pixels, 2188 tiles of 128, the last one partial):
    pixels at the start, around pixel 262144 and at the end, whose 3x3
    patches cross image rows and edges."""
    from unblending_amd import mlp
    L = 3
    params = _perturbed_params(L, 2)
    model = mlp.WarmStartModel(L, params=params)
    w, h = 700, 400
    rgb = np.random.default_rng(1).uniform(0, 1, (h, w, 3))
    print("case forward 700x400 L=3", flush=True)
    y = model.predict(_dev(rgb)).cpu().numpy().reshape(L, -1, 4)
    idx = np.r_[0:64, 262144 - 64:262144 + 64, w * h - 64:w * h]
    X = M.patches(rgb[None], idx, w, h)
    _, _, yo = M.forward(params, X, bf16_ops=True)
    ref = yo[:, :4 * L].reshape(len(idx), L, 4).transpose(1, 0, 2)
    d = np.abs(y[:, idx, :] - ref)
    print("chunked forward max|diff| %.3e" % d.max(), flush=True)
    assert d.max() <= FORWARD_TOL


def test_predict_needs_no_workspace():
    """The C entry point with a null workspace of size 0 (the documented
    size) gives the prediction bit for bit."""
    import ctypes
    import torch
    import unblending_amd as U
    from unblending_amd import mlp
    L = 4
    model = mlp.WarmStartModel(L, params=_perturbed_params(L, 3))
    rgb = _dev(np.random.default_rng(3).uniform(0, 1, (19, 31, 3)))
    print("case predict without workspace 31x19 L=4", flush=True)
    ref = model.predict(rgb)
    out = torch.empty_like(ref)
    lib = U.lib()
    assert lib.unblending_hip_mlp_predict_workspace_size() == 0
    st = lib.unblending_hip_mlp_predict(ctypes.c_void_p(model.data_ptr()), ctypes.c_void_p(rgb.data_ptr()), 31, 19,
                                        L, ctypes.c_void_p(out.data_ptr()), None, 0, None)
    assert st == U.OK
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


def _train_problem(L=4, count=2, w=29, h=19, seed=0):
    rng = np.random.default_rng(seed)
    images = rng.uniform(0, 1, (count, h, w, 3))
    targets = rng.uniform(0.05, 0.95, (count, L, h, w, 4))
    return images, targets


def _check_gradient(g, ref):
    for l, (i, o) in enumerate(M.SHAPES):
        for name, a, b in (("W%d" % (l + 1), g[M.OFF_W[l]:M.OFF_W[l] + i * o], ref[M.OFF_W[l]:M.OFF_W[l] + i * o]),
                           ("b%d" % (l + 1), g[M.OFF_B[l]:M.OFF_B[l] + o], ref[M.OFF_B[l]:M.OFF_B[l] + o])):
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            print("  %s rel %.2e" % (name, rel), flush=True)
            assert rel <= GRAD_REL_TOL, name


def test_gradient_matches_oracle():
    import torch
    from unblending_amd import mlp
    L, B = 4, 256
    images, targets = _train_problem(L)
    count, h, w, _ = images.shape
    params = _perturbed_params(L, 3)
    model = mlp.WarmStartModel(L, params=params)
    idx = np.random.default_rng(4).integers(0, count * h * w, B)
    print("case gradient fixed batch B=256 L=4", flush=True)
    losses, g = model.train(_dev(images), _dev(targets), 1, batch=B, update=False, gradients=True,
                            fixed_samples=torch.from_numpy(idx).cuda())
    X = M.patches(images, idx, w, h)
    T = M.targets_of(targets, idx, L, w, h, 4 * L)
    Lref, gref = M.loss_grad(params, X, T, 4 * L, bf16_ops=True)
    print("loss gpu %.6e oracle %.6e" % (losses[0], Lref), flush=True)
    assert losses[0] == pytest.approx(Lref, rel=LOSS_REL_TOL)
    _check_gradient(g.astype(np.float64), gref)
    assert np.array_equal(model.params(), params)  # update=False leaves the weights


def test_sampled_batch_matches_oracle_sampler():
    from unblending_amd import mlp
    L, B, seed = 3, 128, 0x1234
    images, targets = _train_problem(L, count=3, w=17, h=13, seed=5)
    count, h, w, _ = images.shape
    params = _perturbed_params(L, 6)
    model = mlp.WarmStartModel(L, params=params)
    print("case sampled batch B=128", flush=True)
    losses, g = model.train(_dev(images), _dev(targets), 1, batch=B, seed=seed, update=False, gradients=True)
    idx = M.sample_indices(seed, 0, B, count * h * w)
    X = M.patches(images, idx, w, h)
    T = M.targets_of(targets, idx, L, w, h, 4 * L)
    Lref, gref = M.loss_grad(params, X, T, 4 * L, bf16_ops=True)
    assert losses[0] == pytest.approx(Lref, rel=LOSS_REL_TOL)
    _check_gradient(g.astype(np.float64), gref)


def test_adam_update_matches_oracle():
    import torch
    from unblending_amd import mlp
    L, B, lr = 2, 64, 3e-3
    images, targets = _train_problem(L, seed=7)
    count, h, w, _ = images.shape
    p0 = _perturbed_params(L, 8)
    model = mlp.WarmStartModel(L, params=p0)
    idx = torch.from_numpy(np.random.default_rng(9).integers(0, count * h * w, B)).cuda()
    print("case adam two steps", flush=True)
    m = np.zeros(p0.size, dtype=np.float32)
    v = np.zeros(p0.size, dtype=np.float32)
    p = p0.copy()
    for step in range(2):
        _, g = model.train(_dev(images), _dev(targets), 1, batch=B, lr=lr, fixed_samples=idx, gradients=True)
        p, m, v = M.adam(p.astype(np.float32), m, v, g.astype(np.float32), step, np.float32(lr),
                         np.float32(0.9), np.float32(0.999), np.float32(1e-8))
        got = model.params()
        d = np.abs(got - p)
        print("  step %d max|diff| %.3e" % (step, d.max()), flush=True)
        assert d.max() <= ADAM_TOL
        p = got
    pads = M.unpack(got)[0]
    assert np.all(pads[0][M.INPUTS:, :] == 0.0) and np.all(pads[3][:, 4 * L:] == 0.0)


def _training_set(k, size, seeds):
    from unblending_amd import synth
    return [synth.config(k, size, image_seed=s) for s in seeds]


def _targets(wls):
    import torch
    import unblending_amd as U
    from unblending_amd import device
    h, w, _ = wls[0].rgb.shape
    L = len(wls[0].layers)
    images = torch.stack([_dev(wl.rgb) for wl in wls]).contiguous()
    targets = torch.empty((len(wls), L, h, w, 4), dtype=torch.float64, device="cuda")
    for j, wl in enumerate(wls):
        device.unmix(images[j], layer_infos(wl.layers), out=targets[j])
    return images, targets


def test_training_reduces_loss():
    from unblending_amd import mlp
    wls = _training_set(1, (48, 40), [11, 12])
    print("case training c1 48x40 x2", flush=True)
    images, targets = _targets(wls)
    model = mlp.WarmStartModel(3, seed=1)
    losses = model.fit(images, targets, steps=600, batch=1024, lr=3e-3, chunk=200)
    first, last = float(np.mean(losses[:20])), float(np.mean(losses[-50:]))
    print("loss %.4e -> %.4e" % (first, last), flush=True)
    assert np.all(np.isfinite(losses))
    assert last < 0.3 * first


@pytest.fixture(scope="module")
def c1_model():
    from unblending_amd import mlp
    wls = _training_set(1, (64, 48), [21, 22, 23])
    images, targets = _targets(wls)
    model = mlp.WarmStartModel(3, seed=2)
    model.fit(images, targets, steps=800, batch=2048, lr=3e-3, chunk=200)
    return model


def test_warm_unmix_matches_oracle(oracle, c1_model):
    import unblending_amd as U
    from unblending_amd import device, synth
    wl = synth.config(1, (40, 30), image_seed=99)
    L = len(wl.layers)
    print("case warm unmix c1 40x30", flush=True)
    rgb = _dev(wl.rgb)
    pred = c1_model.predict(rgb).cpu().numpy()
    stats = U.HipStats()
    out = device.unmix(rgb, layer_infos(wl.layers), params=U.hip_params(warm_start=c1_model),
                       stats=stats).cpu().numpy()
    ol = oracle_layers(wl.layers)
    h, w, _ = wl.rgb.shape
    ref = np.empty_like(out)
    capped = 0
    warm_settings = oracle.default_settings()
    warm_settings.max_count = WARM_OUTER_CAP  # the warm attempt's outer cap
    for y in range(h):
        for x in range(w):
            x0 = np.concatenate([pred[:, y, x, 3], pred[:, y, x, :3].reshape(-1)])
            sol, _, cap = oracle.solve_pixel_warm(wl.rgb[y, x], ol, x0, settings=warm_settings)
            if cap:
                capped += 1
                sol, _ = oracle.solve_pixel(wl.rgb[y, x], ol)
            ref[:, y, x, :3] = sol[L:].reshape(L, 3)
            ref[:, y, x, 3] = sol[:L]
    d = np.abs(out - ref)
    exact = float(np.mean(np.all(d == 0, axis=(0, 3))))
    print("warm unmix max|diff| %.3e bit-identical pixels %.4f fallback %d (oracle %d)"
          % (d.max(), exact, stats.fallback_pixels, capped), flush=True)
    assert d.max() <= ALPHA_COLOR_TOL
    assert d.max() == 0.0   # same arithmetic, same order: bit for bit
    assert stats.fallback_pixels == capped
    assert stats.pixels == w * h


def test_warm_start_saves_evaluations(c1_model):
    import unblending_amd as U
    from unblending_amd import device, synth
    wl = synth.config(1, (64, 48), image_seed=77)
    rgb = _dev(wl.rgb)
    infos = layer_infos(wl.layers)
    print("case warm evals c1 64x48", flush=True)
    s_plain, s_warm = U.HipStats(), U.HipStats()
    plain = device.unmix(rgb, infos, stats=s_plain).cpu().numpy()
    warm = device.unmix(rgb, infos, params=U.hip_params(warm_start=c1_model), stats=s_warm).cpu().numpy()
    ratio = s_warm.evals / s_plain.evals
    print("evals plain %d warm %d (ratio %.3f), fallback %d of %d pixels"
          % (s_plain.evals, s_warm.evals, ratio, s_warm.fallback_pixels, plain.shape[1] * plain.shape[2]),
          flush=True)
    # both are points the solver accepts: the recomposites agree with the image alike
    rp, _ = synth.composite(plain[..., 3], plain[..., :3], wl.layers)
    rw, _ = synth.composite(warm[..., 3], warm[..., :3], wl.layers)
    ep, ew = np.sqrt(np.mean((rp - wl.rgb) ** 2)), np.sqrt(np.mean((rw - wl.rgb) ** 2))
    print("recomposite rms vs image: plain %.3e warm %.3e" % (ep, ew), flush=True)
    assert ew <= 1.5 * ep + 1e-4
    assert ratio < 0.9


@pytest.fixture(scope="module")
def c4_model():
    """The warm-start configuration c4: c2's 5-layer set (Normal, Multiply,
    Screen, ColorDodge, Overlay) and generator, a head of 4 L = 20 outputs in
    the network's 32-wide output tile, trained on images of that generator."""
    from unblending_amd import mlp
    wls = _training_set(4, (64, 48), [41, 42, 43])
    assert len(wls[0].layers) == 5
    images, targets = _targets(wls)
    model = mlp.WarmStartModel(5, seed=4)
    model.fit(images, targets, steps=800, batch=2048, lr=3e-3, chunk=200)
    return model


def test_c4_warm_unmix_matches_oracle(oracle, c4_model):
    """c4 at a reduced size: the warm-started first pass equals the oracle's
    warm solve from the same predictions (pixels whose warm attempt hits its
    outer cap solved again from the default start), with the same fallback
    count, and the warm start saves evaluations over the plain solve."""
    import unblending_amd as U
    from unblending_amd import device, synth
    wl = synth.config(4, (36, 28), image_seed=0x5EED0004)
    L = len(wl.layers)
    print("case warm unmix c4 36x28 L=%d" % L, flush=True)
    rgb = _dev(wl.rgb)
    infos = layer_infos(wl.layers)
    pred = c4_model.predict(rgb).cpu().numpy()
    assert pred.shape == (L, 28, 36, 4)
    stats, plain_stats = U.HipStats(), U.HipStats()
    out = device.unmix(rgb, infos, params=U.hip_params(warm_start=c4_model), stats=stats).cpu().numpy()
    device.unmix(rgb, infos, stats=plain_stats)
    ol = oracle_layers(wl.layers)
    h, w, _ = wl.rgb.shape
    ref = np.empty_like(out)
    capped = 0
    warm_settings = oracle.default_settings()
    warm_settings.max_count = WARM_OUTER_CAP
    for y in range(h):
        for x in range(w):
            x0 = np.concatenate([pred[:, y, x, 3], pred[:, y, x, :3].reshape(-1)])
            sol, _, cap = oracle.solve_pixel_warm(wl.rgb[y, x], ol, x0, settings=warm_settings)
            if cap:
                capped += 1
                sol, _ = oracle.solve_pixel(wl.rgb[y, x], ol)
            ref[:, y, x, :3] = sol[L:].reshape(L, 3)
            ref[:, y, x, 3] = sol[:L]
    d = np.abs(out - ref)
    print("c4 warm unmix max|diff| %.3e fallback %d (oracle %d) evals warm %d plain %d"
          % (d.max(), stats.fallback_pixels, capped, stats.evals, plain_stats.evals), flush=True)
    exact = float(np.mean(np.all(d == 0, axis=(0, 3))))
    print("bit-identical pixels %.4f" % exact, flush=True)
    assert d.max() <= ALPHA_COLOR_TOL
    assert d.max() == 0.0   # same arithmetic, same order: bit for bit
    assert stats.fallback_pixels == capped
    assert stats.pixels == w * h
    assert stats.evals < plain_stats.evals
