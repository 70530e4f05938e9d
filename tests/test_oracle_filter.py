"""Pins the oracle's box and guided filters.

Box (image_processing.cpp:120-151, image_processing.hpp:215-222): the
reference's clamp-edge mean over (2r+1)^2 samples is exactly
scipy.ndimage.uniform_filter(mode="nearest"); the O(r^2) brute force and the
sliding-window variant must both equal it.  Guided filter
(image_processing.cpp:214-273): an independent textbook implementation of He
et al.'s colour guided filter (numpy + scipy, per-pixel np.linalg.solve),
plus the closed-form special cases r = 0 (identity) and constant input.
Refinement targets (unblending.cpp:351-399, 311-333): crop, source-over keep,
Plus normalisation, zeros otherwise.
"""
import numpy as np
import pytest
from scipy import ndimage


def box_ref(img, r):
    if r == 0:
        return img.copy()
    return ndimage.uniform_filter(img, size=2 * r + 1, mode="nearest")


@pytest.mark.parametrize("w,h,r", [(7, 5, 0), (7, 5, 1), (7, 5, 2), (7, 5, 3), (7, 5, 8), (16, 9, 4), (1, 6, 2),
                                   (33, 17, 5)])
def test_box_brute_and_sliding_match_clamp_edge_mean(oracle, w, h, r):
    rng = np.random.default_rng(w * 100 + h * 10 + r)
    img = rng.uniform(-1, 2, (h, w))
    ref = box_ref(img, r)
    assert np.abs(oracle.box(img, r, True) - ref).max() < 1e-13
    assert np.abs(oracle.box(img, r, False) - ref).max() < 1e-13


def guided_ref(p, I, r, eps):
    """He et al., colour guide, textbook form."""
    h, w = p.shape
    mI = np.stack([box_ref(I[..., c], r) for c in range(3)], -1)
    mp = box_ref(p, r)
    mIp = np.stack([box_ref(I[..., c] * p, r) for c in range(3)], -1)
    cov = mIp - mI * mp[..., None]
    S = np.empty((h, w, 3, 3))
    for u in range(3):
        for v in range(3):
            S[..., u, v] = box_ref(I[..., u] * I[..., v], r) - mI[..., u] * mI[..., v]
    a = np.linalg.solve(S + eps * np.eye(3), cov[..., None])[..., 0]
    b = mp - np.sum(a * mI, -1)
    return box_ref(b, r) + sum(box_ref(a[..., c], r) * I[..., c] for c in range(3))


@pytest.mark.parametrize("w,h,r,eps", [(7, 5, 1, 1e-4), (12, 10, 2, 1e-4), (9, 14, 3, 1e-3), (6, 4, 8, 1e-2),
                                       (20, 20, 4, 1e-4)])
@pytest.mark.parametrize("brute", [True, False])
def test_guided_filter_textbook(oracle, w, h, r, eps, brute):
    rng = np.random.default_rng(w + h + r)
    I = rng.uniform(0, 1, (h, w, 3))
    p = rng.uniform(0, 1, (h, w))
    q = oracle.guided_filter(p, I, r, eps, brute=brute)
    assert np.abs(q - guided_ref(p, I, r, eps)).max() < 1e-9


def test_guided_filter_special_cases(oracle):
    rng = np.random.default_rng(3)
    I = rng.uniform(0, 1, (11, 13, 3))
    p = rng.uniform(0, 1, (11, 13))
    assert np.array_equal(oracle.guided_filter(p, I, 0, 1e-4, brute=True), p)
    assert np.array_equal(oracle.guided_filter(p, I, 0, 1e-4, brute=False), p)
    c = np.full((11, 13), 0.37)
    assert np.abs(oracle.guided_filter(c, I, 3, 1e-4) - 0.37).max() < 1e-9
    # brute (reference summation) and sliding agree to rounding
    q1 = oracle.guided_filter(p, I, 4, 1e-4, brute=True)
    q2 = oracle.guided_filter(p, I, 4, 1e-4, brute=False)
    assert np.abs(q1 - q2).max() < 1e-10


def test_default_radius(oracle):
    assert oracle.default_radius(1000, 2000) == 60
    assert oracle.default_radius(4096, 4096) == 245
    assert oracle.default_radius(16, 9) == 0
    assert oracle.default_radius(120, 120) == 7


def _first_pass(L, h, w, rng, alpha0=1.0):
    first = rng.uniform(0, 1, (L, h, w, 4))
    first[0, ..., 3] = alpha0
    return first


def test_refine_targets_source_over_keeps_alphas(oracle):
    rng = np.random.default_rng(4)
    L, h, w = 3, 9, 8
    I = rng.uniform(0, 1, (h, w, 3))
    first = _first_pass(L, h, w, rng)
    ls = [oracle.Layer(0, (0.5, 0.5, 0.5), tuple(np.eye(3).ravel() * 50))] * L
    t, bg = oracle.refine_targets(I, ls, first, radius=2, eps=1e-4)
    for l in range(L):
        q = np.clip(oracle.guided_filter(np.ascontiguousarray(first[l, ..., 3]), I, 2, 1e-4), 0, 1)
        assert np.abs(t[l] - q).max() == 0
    assert bg is None


def test_refine_targets_plus_normalises_and_nonopaque_zeroes(oracle):
    rng = np.random.default_rng(5)
    L, h, w = 3, 7, 6
    I = rng.uniform(0, 1, (h, w, 3))
    first = _first_pass(L, h, w, rng, alpha0=0.5)
    plus = [oracle.Layer(0, (0.5, 0.5, 0.5), tuple(np.eye(3).ravel() * 50), (2, 1, 1))] * L
    t, _ = oracle.refine_targets(I, plus, first, radius=1, eps=1e-4)
    assert np.allclose(t.sum(0), 1.0, atol=1e-14)
    so = [oracle.Layer(0, (0.5, 0.5, 0.5), tuple(np.eye(3).ravel() * 50))] * L
    t, _ = oracle.refine_targets(I, so, first, radius=1, eps=1e-4)
    assert np.all(t == 0.0)  # source-over without an opaque background (unblending.cpp:331)


def test_refine_targets_smooth_background(oracle):
    rng = np.random.default_rng(6)
    L, h, w = 2, 8, 9
    I = rng.uniform(0, 1, (h, w, 3))
    first = _first_pass(L, h, w, rng)
    ls = [oracle.Layer(0, (0.5, 0.5, 0.5), tuple(np.eye(3).ravel() * 50))] * L
    _, bg = oracle.refine_targets(I, ls, first, smooth=True, radius=2, eps=1e-4)
    for c in range(3):
        q = np.clip(oracle.guided_filter(np.ascontiguousarray(first[0, ..., c]), I, 2, 1e-4), 0, 1)
        assert np.abs(bg[c] - q).max() == 0
