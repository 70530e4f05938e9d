"""Pins the oracle's per-pixel mathematics.

blend / blend_grad_s / blend_grad_d  (blend_mode.hpp:88-202): hand-computed
  known answers on every branch, central differences away from kinks.
composite (equations.cpp:10-74): hand-computed values and invariants.
energy, constraints (equations.cpp:88-155): closed forms in numpy.
Jacobian (equations.cpp:198-401): central differences of the constraints.
objective gradient (unblending.cpp:67-114): equals dE + J (rho C - lambda)
  with J from the reference-form assembly; central differences for every
  variable in refinement and for the colour block in the first pass, whose
  alpha block is the reference's overwrite (equations.cpp:189-193).
"""
import numpy as np
import pytest

N, MU, SC, OV, DK, LT, CD, CB, HL, SL, DF, EX, LD = range(13)

KAT = [
    # mode, s, d, blend, d/ds, d/dd
    (N, 0.3, 0.7, 0.3, 1.0, 0.0),
    (MU, 0.3, 0.7, 0.21, 0.7, 0.3),
    (SC, 0.3, 0.7, 0.79, 0.3, 0.7),
    (OV, 0.3, 0.4, 0.24, 0.8, 0.6),
    (OV, 0.3, 0.7, 0.58, 0.6, 1.4),
    (DK, 0.3, 0.7, 0.3, 1.0, 0.0),
    (DK, 0.8, 0.7, 0.7, 0.0, 1.0),
    (LT, 0.3, 0.7, 0.7, 0.0, 1.0),
    (LT, 0.8, 0.7, 0.8, 1.0, 0.0),
    (CD, 0.5, 0.25, 0.5, 1.0, 2.0),
    (CD, 0.5, 0.75, 1.0, 0.0, 0.0),
    (CD, 0.5, 0.0, 0.0, 0.0, 0.0),
    (CD, 1.0, 0.5, 1.0, 0.0, 0.0),
    (CB, 0.5, 0.75, 0.5, 1.0, 2.0),
    (CB, 0.5, 0.25, 0.0, 0.0, 0.0),
    (CB, 0.5, 1.0, 1.0, 0.0, 0.0),
    (CB, 0.0, 0.5, 0.0, 0.0, 0.0),
    (HL, 0.25, 0.6, 0.3, 1.2, 0.5),
    (HL, 0.75, 0.6, 0.8, 0.8, 0.5),
    (SL, 0.25, 0.6, 0.48, 0.48, 1.1),
    (SL, 0.75, 0.2, 0.324, 0.496, 1.06),
    (SL, 0.75, 0.64, 0.72, 0.32, 0.8125),
    (DF, 0.3, 0.7, 0.4, -1.0, 1.0),
    (DF, 0.7, 0.3, 0.4, 1.0, -1.0),
    (EX, 0.3, 0.7, 0.58, -0.4, 0.4),
    (LD, 0.3, 0.5, 0.8, 1.0, 1.0),
]


@pytest.mark.parametrize("mode,s,d,b,gs,gd", KAT)
def test_blend_known_answers(oracle, mode, s, d, b, gs, gd):
    assert oracle.blend(s, d, mode) == pytest.approx(b, abs=1e-14)
    assert oracle.blend_grad_s(s, d, mode) == pytest.approx(gs, abs=1e-14)
    assert oracle.blend_grad_d(s, d, mode) == pytest.approx(gd, abs=1e-14)


@pytest.mark.parametrize("mode", range(13))
def test_blend_derivatives_finite_differences(oracle, mode):
    rng = np.random.default_rng(100 + mode)
    h = 1e-6
    checked = 0
    for _ in range(200):
        s, d = rng.uniform(0.02, 0.98, 2)
        # stay away from the branch switches of every mode
        kinks = [0.5, 0.25, d, 1.0 - d]
        if min(abs(s - k) for k in kinks) < 1e-3 or min(abs(d - k) for k in (0.5, 0.25, s)) < 1e-3:
            continue
        if mode == CD and abs(d / (1 - s) - 1) < 1e-3:
            continue
        if mode == CB and abs((1 - d) / s - 1) < 1e-3:
            continue
        fs = (oracle.blend(s + h, d, mode) - oracle.blend(s - h, d, mode)) / (2 * h)
        fd = (oracle.blend(s, d + h, mode) - oracle.blend(s, d - h, mode)) / (2 * h)
        assert oracle.blend_grad_s(s, d, mode) == pytest.approx(fs, abs=1e-6)
        assert oracle.blend_grad_d(s, d, mode) == pytest.approx(fd, abs=1e-6)
        checked += 1
    assert checked > 100


def layers_of(oracle, modes, seed=0, comp=(1, 1, 1)):
    rng = np.random.default_rng(seed)
    out = []
    for m in modes:
        A = rng.standard_normal((3, 3))
        P = A @ A.T + 3 * np.eye(3)
        out.append(oracle.Layer(m, tuple(rng.uniform(0.1, 0.9, 3)), tuple(P.ravel()), comp))
    return out


def pack(alphas, colors):
    return np.concatenate([np.asarray(alphas, float), np.asarray(colors, float).ravel()])


def test_composite_two_layer_multiply(oracle):
    ls = layers_of(oracle, [N, MU])
    x = pack([1.0, 0.5], [[0.5, 0.5, 0.5], [0.4, 0.6, 0.8]])
    assert np.allclose(oracle.composite(x, ls), [0.35, 0.4, 0.45, 1.0], atol=1e-15)


def test_composite_invariants(oracle):
    rng = np.random.default_rng(1)
    for L in range(1, 8):
        ls = layers_of(oracle, rng.integers(0, 13, L), seed=L)
        for _ in range(20):
            a = rng.uniform(0, 1, L)
            a[0] = 1.0
            c = rng.uniform(0, 1, (L, 3))
            out = oracle.composite(pack(a, c), ls)
            assert out[3] == pytest.approx(1.0, abs=1e-15)  # opaque source-over stack stays opaque
    # a fully opaque Normal top layer hides everything below
    ls = layers_of(oracle, [MU, SC, N])
    x = pack([1.0, 0.3, 1.0], [[0.1, 0.2, 0.3], [0.6, 0.5, 0.4], [0.9, 0.8, 0.7]])
    assert np.allclose(oracle.composite(x, ls)[:3], [0.9, 0.8, 0.7], atol=1e-15)
    # Normal over opaque: linear interpolation
    ls = layers_of(oracle, [N, N])
    x = pack([1.0, 0.25], [[0.2, 0.4, 0.6], [1.0, 0.0, 0.5]])
    assert np.allclose(oracle.composite(x, ls)[:3], 0.25 * np.array([1.0, 0.0, 0.5]) + 0.75 * np.array([0.2, 0.4, 0.6]),
                       atol=1e-15)
    # Plus: A = a_s + a_d
    ls = layers_of(oracle, [N, N], comp=(2, 1, 1))
    x = pack([0.3, 0.4], [[0.2, 0.4, 0.6], [1.0, 0.0, 0.5]])
    assert oracle.composite(x, ls)[3] == pytest.approx(0.7, abs=1e-15)


def energy_np(x, ls, minimum):
    L = len(ls)
    e = 0.0
    for i in range(L):
        d = x[L + 3 * i:L + 3 * i + 3] - np.array(ls[i].mean)
        e += x[i] * d @ np.array(ls[i].precision).reshape(3, 3) @ d
    return e + (0.01 * x[:L].sum() if minimum else 0.0)


def random_point(rng, L):
    a = rng.uniform(0.05, 0.95, L)
    a[0] = 1.0
    return pack(a, rng.uniform(0.05, 0.95, (L, 3)))


@pytest.mark.parametrize("L", [1, 2, 3, 5, 7])
def test_energy_and_constraints_closed_form(oracle, L):
    rng = np.random.default_rng(L)
    ls = layers_of(oracle, rng.integers(0, 13, L), seed=L)
    for _ in range(10):
        x = random_point(rng, L)
        for minimum in (0, 1):
            assert oracle.energy(x, ls, minimum) == pytest.approx(energy_np(x, ls, minimum), rel=1e-13)
        t = rng.uniform(0, 1, 3)
        comp = oracle.composite(x, ls)
        C = oracle.constraints(x, ls, t)
        assert np.allclose(C, np.concatenate([comp[:3] - t, [comp[3] - 1.0]]), atol=0)
        ta = rng.uniform(0, 1, L)
        C = oracle.constraints(x, ls, t, ta)
        assert np.allclose(C, np.concatenate([comp[:3] - t, x[:L] - ta]), atol=0)


def central_jacobian(fn, x, h=1e-6):
    cols = []
    for v in range(x.size):
        e = np.zeros_like(x)
        e[v] = h
        cols.append((fn(x + e) - fn(x - e)) / (2 * h))
    return np.array(cols)  # [var][constraint]


@pytest.mark.parametrize("L", [2, 3, 5, 7])
@pytest.mark.parametrize("refine", [False, True])
def test_constraint_jacobian_finite_differences(oracle, L, refine):
    rng = np.random.default_rng(10 * L + refine)
    modes = [N, MU, SC, OV, HL, SL, EX][:L]
    ls = layers_of(oracle, modes, seed=L)
    t = rng.uniform(0, 1, 3)
    ta = rng.uniform(0, 1, L) if refine else None
    for _ in range(5):
        x = random_point(rng, L)
        x[0] = 0.9  # exercise the alpha_0 column as well
        J = oracle.constraint_jacobian(x, ls, refine)
        Jfd = central_jacobian(lambda z: oracle.constraints(z, ls, t, ta), x)
        assert np.abs(J - Jfd).max() < 1e-6


@pytest.mark.parametrize("L", [1, 2, 3, 5, 7])
def test_objective_gradient_equals_reference_assembly(oracle, L):
    """Adjoint sweep == dE + J (rho C - lambda) with the reference's J."""
    rng = np.random.default_rng(20 + L)
    ls = layers_of(oracle, rng.integers(0, 13, L), seed=L + 50)
    for refine in (False, True):
        for _ in range(10):
            x = random_point(rng, L)
            x[0] = rng.uniform(0.5, 1.0)
            t = rng.uniform(0, 1, 3)
            ta = rng.uniform(0, 1, L) if refine else None
            nc = 3 + L if refine else 4
            lam = rng.standard_normal(nc)
            rho = float(rng.uniform(1, 1e4))
            f, g = oracle.objective(x, ls, t, ta, lam, rho)
            C = oracle.constraints(x, ls, t, ta)
            J = oracle.constraint_jacobian(x, ls, refine)
            dE = oracle.energy_gradient(x, ls, not refine)
            ref = dE + J @ (rho * C - lam)
            assert np.abs(g - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())
            e = energy_np(x, ls, not refine)
            assert f == pytest.approx(e - lam @ C + 0.5 * rho * C @ C, rel=1e-12)


@pytest.mark.parametrize("L", [2, 3, 5, 7])
def test_objective_gradient_finite_differences(oracle, L):
    rng = np.random.default_rng(30 + L)
    modes = [N, MU, SC, OV, HL, SL, EX][:L]
    ls = layers_of(oracle, modes, seed=L + 70)
    for refine in (False, True):
        x = random_point(rng, L)
        t = rng.uniform(0, 1, 3)
        ta = rng.uniform(0, 1, L) if refine else None
        nc = 3 + L if refine else 4
        lam = rng.standard_normal(nc)
        rho = 37.0
        _, g = oracle.objective(x, ls, t, ta, lam, rho)
        fd = central_jacobian(lambda z: np.array([oracle.objective(z, ls, t, ta, lam, rho)[0]]), x)[:, 0]
        if refine:
            assert np.abs(g - fd).max() < 1e-5
        else:
            assert np.abs(g[L:] - fd[L:]).max() < 1e-5
            # alpha block: the reference assigns 0.01 in place of D(c_i) (equations.cpp:189-193)
            dist = np.array([energy_np(np.concatenate([[1.0] + [0.0] * (L - 1) if i == 0 else
                                                       [0.0] * i + [1.0] + [0.0] * (L - 1 - i), x[L:]]), ls, False)
                             for i in range(L)])
            assert np.abs(g[:L] - (fd[:L] - dist)).max() < 1e-5


def test_energy_gradient_alpha_block_quirk(oracle):
    rng = np.random.default_rng(5)
    ls = layers_of(oracle, [N, MU, SC], seed=9)
    x = random_point(rng, 3)
    g = oracle.energy_gradient(x, ls, True)
    assert np.all(g[:3] == 0.01)
    g = oracle.energy_gradient(x, ls, False)
    for i in range(3):
        d = x[3 + 3 * i:6 + 3 * i] - np.array(ls[i].mean)
        assert g[i] == pytest.approx(d @ np.array(ls[i].precision).reshape(3, 3) @ d, rel=1e-13)
