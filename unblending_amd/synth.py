"""Seeded synthetic workloads for the two passes.

Each workload is an RGB image made by compositing randomly coloured layers
bottom to top with the blend modes of a layer set, plus the layer set itself
(one Gaussian colour model a layer).  The recipes are the benchmark
configurations of the unmixing + refinement path:

  config 1   256 x 256, Normal / Multiply / Screen, Sigma = 0.01 I, alphas of
             blurred noise (sigma 8 px), noise 0.005, filter r = 8, eps = 1e-4
  config 2   1024 x 1024, Normal / Multiply / Screen / ColorDodge / Overlay,
             full SPD covariances (eigenvalues log-uniform [0.002, 0.02]),
             blurred noise (sigma 16 px) with 30 % of pixels set to 0,
             noise 0.005, r = 16, eps = 1e-4
  config 3   4096 x 4096, Normal / Multiply / Screen / Overlay / ColorBurn /
             HardLight / SoftLight, means in [0.05, 0.95], eigenvalues
             log-uniform [0.001, 0.03], alphas = 20 Gaussian blobs
             (sigma 20-200 px), noise 0.01, r = 32, eps = 1e-3
  config 5   64 images, sides log-uniform [256, 2048], aspect [0.5, 2], 3-7
             layers each from ten modes, config 2's models, r = side / 64

The bottom layer is Normal and opaque; every comp-op is source-over.  The
seed of config k is 0x5EED0000 + k (image j of config 5: 0x5EED0500 + j).
`size` shrinks an image for tests while keeping the recipe; `image_seed`
keeps config k's layer set and draws another image of it (the alpha fields
and noise come from that seed).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy import ndimage

NORMAL, MULTIPLY, SCREEN, OVERLAY, DARKEN, LIGHTEN, COLOR_DODGE, COLOR_BURN, HARD_LIGHT, \
    SOFT_LIGHT, DIFFERENCE, EXCLUSION, LINEAR_DODGE = range(13)

MODE_NAMES = ["Normal", "Multiply", "Screen", "Overlay", "Darken", "Lighten", "ColorDodge",
              "ColorBurn", "HardLight", "SoftLight", "Difference", "Exclusion", "LinearDodge"]


@dataclass
class LayerModel:
    mode: int
    mean: np.ndarray        # (3,)
    precision: np.ndarray   # (3, 3) symmetric positive definite
    comp: tuple = (1, 1, 1)
    covariance: np.ndarray = field(default=None, repr=False)


@dataclass
class Workload:
    name: str
    rgb: np.ndarray          # (h, w, 3) float64
    layers: list             # [LayerModel]
    radius: int
    eps: float
    opaque: bool = True
    smooth: bool = False

    @property
    def width(self):
        return self.rgb.shape[1]

    @property
    def height(self):
        return self.rgb.shape[0]


def blend(s, d, mode):
    """Vectorised separable blend (crop = false)."""
    e = 1e-05
    if mode == NORMAL:
        return s
    if mode == MULTIPLY:
        return s * d
    if mode == SCREEN:
        return 1.0 - (1.0 - s) * (1.0 - d)
    if mode == OVERLAY:
        return np.where(d <= 0.5, 2.0 * s * d, 1.0 - 2.0 * (1.0 - s) * (1.0 - d))
    if mode == DARKEN:
        return np.minimum(s, d)
    if mode == LIGHTEN:
        return np.maximum(s, d)
    if mode == COLOR_DODGE:
        with np.errstate(divide="ignore", invalid="ignore"):
            q = np.minimum(1.0, d / (1.0 - s))
        return np.where(d < e, 0.0, np.where(1.0 - s < e, 1.0, q))
    if mode == COLOR_BURN:
        with np.errstate(divide="ignore", invalid="ignore"):
            q = 1.0 - np.minimum(1.0, (1.0 - d) / s)
        return np.where(1.0 - d < e, 1.0, np.where(s < e, 0.0, q))
    if mode == HARD_LIGHT:
        return np.where(s <= 0.5, 2.0 * s * d, 1.0 - 2.0 * (1.0 - s) * (1.0 - d))
    if mode == SOFT_LIGHT:
        g = np.where(d <= 0.25, ((16.0 * d - 12.0) * d + 4.0) * d, np.sqrt(np.maximum(d, 0.0)))
        return np.where(s <= 0.5, d - (1.0 - 2.0 * s) * d * (1.0 - d), d + (2.0 * s - 1.0) * (g - d))
    if mode == DIFFERENCE:
        return np.abs(s - d)
    if mode == EXCLUSION:
        return s + d - 2.0 * s * d
    if mode == LINEAR_DODGE:
        return s + d
    raise ValueError(mode)


def random_rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def spd_model(rng, mode, mean_lo, mean_hi, eig_lo, eig_hi):
    mean = rng.uniform(mean_lo, mean_hi, 3)
    R = random_rotation(rng)
    lam = np.exp(rng.uniform(np.log(eig_lo), np.log(eig_hi), 3))
    cov = R @ np.diag(lam) @ R.T
    cov = 0.5 * (cov + cov.T)
    P = np.linalg.inv(cov)
    P = 0.5 * (P + P.T)
    return LayerModel(mode, mean, P, covariance=cov)


def iso_model(rng, mode, var):
    mean = rng.uniform(0.1, 0.9, 3)
    cov = var * np.eye(3)
    return LayerModel(mode, mean, np.eye(3) / var, covariance=cov)


def blurred_noise(rng, h, w, sigma):
    u = rng.uniform(0.0, 1.0, (h, w))
    b = ndimage.gaussian_filter(u, sigma, mode="nearest", truncate=3.0)
    lo, hi = b.min(), b.max()
    return (b - lo) / (hi - lo) if hi > lo else np.zeros_like(b)


def blobs(rng, h, w, count=20, smin=20.0, smax=200.0):
    out = np.zeros((h, w))
    ys = np.arange(h, dtype=np.float64)
    xs = np.arange(w, dtype=np.float64)
    for _ in range(count):
        cy, cx = rng.uniform(0, h), rng.uniform(0, w)
        s = rng.uniform(smin, smax)
        amp = rng.uniform(0.5, 1.0)
        gy = np.exp(-0.5 * ((ys - cy) / s) ** 2)
        gx = np.exp(-0.5 * ((xs - cx) / s) ** 2)
        out += amp * np.outer(gy, gx)
    return np.clip(out, 0.0, 1.0)


def composite(alphas, colors, layers):
    """Bottom-to-top source-over fold, as the reference's composite_layers."""
    color = colors[0].copy()
    alpha = alphas[0].copy()
    for k in range(1, len(layers)):
        a_s = alphas[k][..., None]
        a_d = alpha[..., None]
        A = a_s + a_d - a_s * a_d
        f = blend(colors[k], color, layers[k].mode)
        pre = f * a_s * a_d + a_s * (1.0 - a_d) * colors[k] + a_d * (1.0 - a_s) * color
        color = np.where(A > 1e-12, pre / np.where(A > 1e-12, A, 1.0), pre)
        alpha = A[..., 0]
    return color, alpha


def _layer_colors(rng, model, h, w):
    L = np.linalg.cholesky(model.covariance)
    z = rng.standard_normal((h, w, 3))
    return np.clip(model.mean + z @ L.T, 0.0, 1.0)


def _render(rng, layers, alpha_fields, noise, h, w):
    color = None
    alpha = None
    for k, model in enumerate(layers):
        c = _layer_colors(rng, model, h, w)
        a = np.ones((h, w)) if k == 0 else alpha_fields[k]
        if k == 0:
            color, alpha = c, a
            continue
        a_s = a[..., None]
        a_d = alpha[..., None]
        A = a_s + a_d - a_s * a_d
        f = blend(c, color, model.mode)
        pre = f * a_s * a_d + a_s * (1.0 - a_d) * c + a_d * (1.0 - a_s) * color
        color = np.where(A > 1e-12, pre / np.where(A > 1e-12, A, 1.0), pre)
        alpha = A[..., 0]
    rgb = color + rng.normal(0.0, noise, color.shape)
    return np.ascontiguousarray(np.clip(rgb, 0.0, 1.0))


def _image_rng(rng, image_seed):
    return rng if image_seed is None else np.random.default_rng(image_seed)


def config1(size=(256, 256), seed=0x5EED0001, image_seed=None):
    rng = np.random.default_rng(seed)
    w, h = size
    modes = [NORMAL, MULTIPLY, SCREEN]
    layers = [iso_model(rng, m, 0.01) for m in modes]
    rng = _image_rng(rng, image_seed)
    fields = [None] + [blurred_noise(rng, h, w, 8.0) for _ in modes[1:]]
    rgb = _render(rng, layers, fields, 0.005, h, w)
    return Workload("c1 %dx%d L=3" % (w, h), rgb, layers, 8, 1e-4)


def config2(size=(1024, 1024), seed=0x5EED0002, image_seed=None):
    rng = np.random.default_rng(seed)
    w, h = size
    modes = [NORMAL, MULTIPLY, SCREEN, COLOR_DODGE, OVERLAY]
    layers = [spd_model(rng, m, 0.1, 0.9, 0.002, 0.02) for m in modes]
    rng = _image_rng(rng, image_seed)
    fields = [None]
    for _ in modes[1:]:
        a = blurred_noise(rng, h, w, 16.0)
        a[rng.uniform(0.0, 1.0, (h, w)) < 0.3] = 0.0
        fields.append(a)
    rgb = _render(rng, layers, fields, 0.005, h, w)
    return Workload("c2 %dx%d L=5" % (w, h), rgb, layers, 16, 1e-4)


def config3(size=(4096, 4096), seed=0x5EED0003, image_seed=None):
    rng = np.random.default_rng(seed)
    w, h = size
    modes = [NORMAL, MULTIPLY, SCREEN, OVERLAY, COLOR_BURN, HARD_LIGHT, SOFT_LIGHT]
    layers = [spd_model(rng, m, 0.05, 0.95, 0.001, 0.03) for m in modes]
    rng = _image_rng(rng, image_seed)
    scale = min(w, h) / 4096.0
    fields = [None] + [blobs(rng, h, w, 20, 20.0 * max(scale, 0.05), 200.0 * max(scale, 0.05))
                       for _ in modes[1:]]
    rgb = _render(rng, layers, fields, 0.01, h, w)
    return Workload("c3 %dx%d L=7" % (w, h), rgb, layers, 32, 1e-3)


def config4(size=(4096, 4096), image_seed=None):
    """The warm-start configuration: config 2's layer set and generator (5
    layers, r = 16, eps = 1e-4) on a fresh seeded image, 4096 x 4096 by
    default; training images pass their own image_seed."""
    wl = config2(size, image_seed=0x5EED0004 if image_seed is None else image_seed)
    w, h = size
    return Workload("c4 %dx%d L=5" % (w, h), wl.rgb, wl.layers, wl.radius, wl.eps)


C5_MODES = [NORMAL, MULTIPLY, SCREEN, OVERLAY, COLOR_DODGE, COLOR_BURN, HARD_LIGHT, SOFT_LIGHT,
            DARKEN, LIGHTEN]


def config5_image(j, scale=1.0, image_seed=None):
    """Image j of the batch configuration; image_seed: another image of the
    same size and layer set (training images of a per-set model)."""
    rng = np.random.default_rng(0x5EED0500 + j)
    side = np.exp(rng.uniform(np.log(256), np.log(2048))) * scale
    aspect = rng.uniform(0.5, 2.0)
    w = max(8, int(round(side * np.sqrt(aspect))))
    h = max(8, int(round(side / np.sqrt(aspect))))
    L = int(rng.integers(3, 8))
    modes = [NORMAL] + [int(m) for m in rng.choice(C5_MODES[1:], L - 1)]
    layers = [spd_model(rng, m, 0.1, 0.9, 0.002, 0.02) for m in modes]
    rng = _image_rng(rng, image_seed)
    fields = [None] + [blurred_noise(rng, h, w, 16.0 * scale) for _ in modes[1:]]
    rgb = _render(rng, layers, fields, 0.005, h, w)
    r = max(1, min(w, h) // 64)
    return Workload("c5[%d] %dx%d L=%d" % (j, w, h, L), rgb, layers, r, 1e-4)


def config(k, size=None, image_seed=None):
    if k == 1:
        return config1(size or (256, 256), image_seed=image_seed)
    if k == 2:
        return config2(size or (1024, 1024), image_seed=image_seed)
    if k == 3:
        return config3(size or (4096, 4096), image_seed=image_seed)
    if k == 4:
        return config4(size or (4096, 4096), image_seed=image_seed)
    raise ValueError("config %r" % k)
