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
        q = np.clip(oracle.guided_filter_exact(np.ascontiguousarray(first[l, ..., 3]), I, 2, 1e-4), 0, 1)
        assert np.abs(t[l] - q).max() == 0
    assert bg is None
    # the FP64-sum filters give the same targets to within the coefficient grids (2^-35 here)
    ts, _ = oracle.refine_targets(I, ls, first, radius=2, eps=1e-4, filter="sliding")
    assert np.abs(ts - t).max() < 1e-10


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
        q = np.clip(oracle.guided_filter_exact(np.ascontiguousarray(first[0, ..., c]), I, 2, 1e-4), 0, 1)
        assert np.abs(bg[c] - q).max() == 0


# ---- the exact-sum filter (the device filter's definition) ----

def _qexp(x):
    """smallest e with |x| < 2^e (zero: -1022), from the IEEE exponent"""
    x = abs(float(x))
    if x == 0.0:
        return -1022
    m, e = np.frexp(x)  # x = m 2^e, m in [0.5, 1)
    return int(e)


def _qprefix(r):
    """longest run of window sums the device adds up for radius r"""
    return 256 if r <= 96 else 1024


def _qscale(e, r):
    n = _qprefix(r) * (2 * r + 1)
    return min(53 - e - int(np.ceil(np.log2(n))), 1000)


def _coeff_bounds(imax, pmax, eps):
    """|a_c| <= pmax / sqrt(eps), |b| <= pmax + s3 imax pmax / sqrt(eps)"""
    ba = pmax * (1.0 / np.sqrt(eps))
    return ba, pmax + (1.7320508075688774 * imax) * ba


def _qbox_brute(v, r, e):
    """window means of quantised values, brute-force window sums in Python
    integers (no sliding, no int64 wrap)"""
    h, w = v.shape
    if r == 0:
        return v.copy()
    k = _qscale(e, r)
    q = [[int(np.rint(np.ldexp(v[y, x], k))) for x in range(w)] for y in range(h)]
    wgt = 1.0 / float((2 * r + 1) ** 2)
    out = np.zeros_like(v)
    for y in range(h):
        for x in range(w):
            s = 0
            for dy in range(-r, r + 1):
                for dx in range(-r, r + 1):
                    s += q[min(max(y + dy, 0), h - 1)][min(max(x + dx, 0), w - 1)]
            assert abs(s) <= 2 ** 53
            out[y, x] = np.ldexp(float(s), -k) * wgt
    return out


def _exact_filter_numpy(p, I, r, eps):
    """Independent restatement of the exact-sum guided filter."""
    imax, pmax = float(np.abs(I).max()), float(np.abs(p).max())
    eI, eP = _qexp(imax), _qexp(pmax)
    eG, eC = max(eI, 2 * eI), max(eP, eI + eP)
    ba, bb = _coeff_bounds(imax, pmax, eps)
    eA, eB = _qexp(ba), _qexp(bb)
    Ic = [I[..., c] for c in range(3)]
    mI = [_qbox_brute(Ic[c], r, eG) for c in range(3)]
    mp = _qbox_brute(p, r, eC)
    cov = [_qbox_brute(Ic[c] * p, r, eC) - mI[c] * mp for c in range(3)]
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    var = [_qbox_brute(Ic[u] * Ic[v], r, eG) - mI[u] * mI[v] for u, v in pairs]
    m00, m01, m02 = var[0] + eps, var[1], var[2]
    m11, m12, m22 = var[3] + eps, var[4], var[5] + eps
    c00 = m11 * m22 - m12 * m12
    c01 = m02 * m12 - m01 * m22
    c02 = m01 * m12 - m02 * m11
    c11 = m00 * m22 - m02 * m02
    c12 = m01 * m02 - m00 * m12
    c22 = m00 * m11 - m01 * m01
    det = m00 * c00 + m01 * c01 + m02 * c02
    inv = 1.0 / det
    a0 = (c00 * inv) * cov[0] + (c01 * inv) * cov[1] + (c02 * inv) * cov[2]
    a1 = (c01 * inv) * cov[0] + (c11 * inv) * cov[1] + (c12 * inv) * cov[2]
    a2 = (c02 * inv) * cov[0] + (c12 * inv) * cov[1] + (c22 * inv) * cov[2]
    b = ((mp - a0 * mI[0]) - a1 * mI[1]) - a2 * mI[2]
    q = _qbox_brute(b, r, eB)
    for c, a in enumerate((a0, a1, a2)):
        q = q + _qbox_brute(a, r, eA) * Ic[c]
    return q


def test_quantisation_exponents(oracle):
    for x, e in [(1.0, 1), (0.99, 0), (0.5, 0), (0.4999, -1), (2.0, 2), (3.0, 2), (0.0, -1022), (-1.5, 1)]:
        assert oracle.lib().orc_qexp(x) == e == _qexp(x), x
    assert oracle.lib().orc_qscale(1, 1) == 53 - 1 - 10  # 256 * 3 = 768 <= 2^10
    assert oracle.lib().orc_qscale(0, 32) == 53 - 15  # 256 * 65 = 16640 <= 2^15
    assert oracle.lib().orc_qscale(0, 96) == 53 - 16  # 256 * 193 = 49408 <= 2^16
    assert oracle.lib().orc_qscale(0, 97) == 53 - 18  # 1024 * 195 = 199680 <= 2^18
    assert oracle.lib().orc_qscale(-1022, 3) == 1000
    for r in (1, 8, 32, 96, 97, 480):
        assert oracle.lib().orc_qprefix(r) == _qprefix(r)
        # every partial sum of the device (a prefix of qprefix(r) window sums,
        # each of 2r+1 values of magnitude <= 2^(e+k)) stays within 2^53
        k = oracle.lib().orc_qscale(3, r)
        assert _qprefix(r) * (2 * r + 1) * 2.0 ** (3 + k) <= 2.0 ** 53
    import ctypes
    ba, bb = ctypes.c_double(), ctypes.c_double()
    oracle.lib().orc_coeff_bounds(0.9, 0.7, 1e-3, ctypes.byref(ba), ctypes.byref(bb))
    assert (ba.value, bb.value) == _coeff_bounds(0.9, 0.7, 1e-3)


@pytest.mark.parametrize("w,h,r,eps", [(7, 5, 1, 1e-4), (6, 4, 2, 1e-3), (5, 5, 3, 1e-4), (4, 3, 8, 1e-4)])
def test_exact_filter_matches_independent_restatement(oracle, w, h, r, eps):
    rng = np.random.default_rng(w * 100 + h * 10 + r)
    I = rng.uniform(0, 1, (h, w, 3))
    p = rng.uniform(0, 1, (h, w))
    assert np.array_equal(oracle.guided_filter_exact(p, I, r, eps), _exact_filter_numpy(p, I, r, eps))


@pytest.mark.parametrize("w,h,r,eps", [(23, 17, 3, 1e-4), (40, 9, 12, 1e-3), (16, 16, 5, 1e-4)])
def test_exact_filter_is_the_guided_filter_to_rounding(oracle, w, h, r, eps):
    """Quantisation moves the statistics by at most the grid step, far
    below the filter's 1e-6 bar: the exact and FP64 brute-force filters
    agree to 1e-9."""
    rng = np.random.default_rng(w + h + r)
    I = rng.uniform(0, 1, (h, w, 3))
    p = rng.uniform(0, 1, (h, w))
    qe = oracle.guided_filter_exact(p, I, r, eps)
    qb = oracle.guided_filter(p, I, r, eps, brute=True)
    assert np.abs(qe - qb).max() < 1e-9


@pytest.mark.parametrize("w,h,r", [(3, 200, 100), (200, 3, 100), (70, 5, 130), (2, 2, 480), (5, 9, 300)])
def test_exact_filter_at_large_radii(oracle, w, h, r):
    """Beyond r = 96 the partial sums run over 1024 columns of wider windows,
    so the grids are coarser.  The coarsest is the coefficients' (|a_c| <=
    max|p| / sqrt(eps) = 100 here: a step of 2^-26 at r = 480, half of it
    the rounding error of each of a_r, a_g, a_b, and b's twice that), so
    the exact filter stays within about 4e-8 of the FP64 brute-force one at
    eps = 1e-4 (measured up to 2.2e-8), below the 1e-6 bar."""
    rng = np.random.default_rng(w * 1000 + h + r)
    I = rng.uniform(0, 1, (h, w, 3))
    p = rng.uniform(0, 1, (h, w))
    qe = oracle.guided_filter_exact(p, I, r, 1e-4)
    qb = oracle.guided_filter(p, I, r, 1e-4, brute=True)
    assert np.abs(qe - qb).max() < 5e-8


@pytest.mark.parametrize("eps", [1e-4, 1e-3, 1e-2])
def test_coefficient_bound_holds_with_margin(eps):
    """The grids of a and b assume |a|_2 <= sd(p) / (2 sqrt(eps)) (the
    statistics are covariances of one data set; the bound uses twice that).
    Checked on the textbook filter's coefficients for adversarial guides:
    nearly flat, nearly collinear channels, and p a function of the guide."""
    rng = np.random.default_rng(int(eps * 1e5))
    h, w, r = 24, 20, 3
    base = rng.uniform(0, 1, (h, w))
    cases = [rng.uniform(0, 1, (h, w, 3)),
             np.stack([base, base + 1e-3 * rng.uniform(0, 1, (h, w)), 0.5 * base], -1),
             0.5 + 1e-4 * rng.uniform(0, 1, (h, w, 3))]
    for I in cases:
        for p in (rng.uniform(0, 1, (h, w)), I[..., 0], (I[..., 1] > 0.5).astype(float)):
            mI = np.stack([box_ref(I[..., c], r) for c in range(3)], -1)
            mp = box_ref(p, r)
            cov = np.stack([box_ref(I[..., c] * p, r) for c in range(3)], -1) - mI * mp[..., None]
            S = np.empty((h, w, 3, 3))
            for u in range(3):
                for v in range(3):
                    S[..., u, v] = box_ref(I[..., u] * I[..., v], r) - mI[..., u] * mI[..., v]
            a = np.linalg.solve(S + eps * np.eye(3), cov[..., None])[..., 0]
            sdp = np.sqrt(np.maximum(box_ref(p * p, r) - mp * mp, 0.0))
            assert np.all(np.linalg.norm(a, axis=-1) <= sdp / (2 * np.sqrt(eps)) * (1 + 1e-6) + 1e-9)
            ba, bb = _coeff_bounds(float(np.abs(I).max()), float(np.abs(p).max()), eps)
            b = mp - np.sum(a * mI, -1)
            assert np.abs(a).max() <= 0.5 * ba * (1 + 1e-6) and np.abs(b).max() <= bb


def test_exact_filter_invariants(oracle):
    rng = np.random.default_rng(3)
    I = rng.uniform(0, 1, (11, 13, 3))
    p = rng.uniform(0, 1, (11, 13))
    assert np.array_equal(oracle.guided_filter_exact(p, I, 0, 1e-4), p)  # r = 0: identity
    c = np.full((11, 13), 0.37)
    # within half a step of b's grid: |b| <= 64.4 < 2^7, k = 53 - 7 - 11, step 2^-35
    assert np.abs(oracle.guided_filter_exact(c, I, 3, 1e-4) - 0.37).max() <= 2.0 ** -36
    # bounds passed explicitly equal to the inputs' give the same result; larger bounds only a coarser grid
    q0 = oracle.guided_filter_exact(p, I, 3, 1e-4)
    q1 = oracle.guided_filter_exact(p, I, 3, 1e-4, imax=float(np.abs(I).max()), pmax=float(np.abs(p).max()))
    assert np.array_equal(q0, q1)
    q2 = oracle.guided_filter_exact(p, I, 3, 1e-4, imax=4.0, pmax=4.0)
    assert np.abs(q2 - q0).max() < 1e-8


def test_planted_mistake_in_restatement_is_caught(oracle):
    """A dropped clamp at the left edge (a plausible indexing slip) changes
    the exact filter and fails the bit-exact comparison."""
    rng = np.random.default_rng(9)
    I = rng.uniform(0, 1, (5, 7, 3))
    p = rng.uniform(0, 1, (5, 7))
    ref = oracle.guided_filter_exact(p, I, 1, 1e-4)
    shifted = np.roll(p, 1, axis=1)  # the restatement fed a misaligned plane
    assert not np.array_equal(ref, _exact_filter_numpy(shifted, I, 1, 1e-4))
