"""ctypes binding of the oracle library (oracle/_build/libub_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (unblending_amd/)
never imports it.  See ub_oracle.h for what each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libub_oracle.so")

MAX_LAYERS = 16


class OrcLayer(ctypes.Structure):
    _fields_ = [
        ("comp_x", ctypes.c_int32),
        ("comp_y", ctypes.c_int32),
        ("comp_z", ctypes.c_int32),
        ("blend_mode", ctypes.c_int32),
        ("mean", ctypes.c_double * 3),
        ("precision", ctypes.c_double * 9),
    ]


class OrcSettings(ctypes.Structure):
    _fields_ = [
        ("memory", ctypes.c_int32),
        ("max_evals", ctypes.c_int32),
        ("max_count", ctypes.c_int32),
        ("max_ls", ctypes.c_int32),
        ("ftol_rel", ctypes.c_double),
        ("xtol_rel", ctypes.c_double),
        ("epsilon", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("initial_rho", ctypes.c_double),
        ("alpha_fixed", ctypes.c_int32),
        ("gray_mask", ctypes.c_int32),
        ("sparsity", ctypes.c_int32),
        ("sigma", ctypes.c_double),
    ]


class OrcVariant(ctypes.Structure):
    _fields_ = [("gray_mask", ctypes.c_int32), ("sparsity", ctypes.c_int32), ("sigma", ctypes.c_double)]


class OrcCounters(ctypes.Structure):
    _fields_ = [
        ("evals", ctypes.c_int64),
        ("iterations", ctypes.c_int64),
        ("outer", ctypes.c_int64),
        ("pixels", ctypes.c_int64),
        ("max_pairs", ctypes.c_int64),
        ("pairs_used", ctypes.c_int64),
    ]


ORC_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(OrcLayer)
        L.orc_blend.restype = ctypes.c_double
        L.orc_blend.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_blend_grad_s.restype = ctypes.c_double
        L.orc_blend_grad_s.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_blend_grad_d.restype = ctypes.c_double
        L.orc_blend_grad_d.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_composite_stack.argtypes = [dp, ctypes.c_int, lp, dp]
        L.orc_distance.restype = ctypes.c_double
        L.orc_distance.argtypes = [dp, lp]
        L.orc_energy.restype = ctypes.c_double
        L.orc_energy.argtypes = [dp, ctypes.c_int, lp, ctypes.c_int]
        L.orc_constraints.restype = ctypes.c_int
        L.orc_constraints.argtypes = [dp, ctypes.c_int, lp, dp, dp, dp]
        L.orc_energy_gradient.argtypes = [dp, ctypes.c_int, lp, ctypes.c_int, dp]
        L.orc_constraint_jacobian.restype = ctypes.c_int
        L.orc_constraint_jacobian.argtypes = [dp, ctypes.c_int, lp, ctypes.c_int, dp]
        L.orc_objective.restype = ctypes.c_double
        L.orc_objective.argtypes = [dp, ctypes.c_int, lp, dp, dp, dp, ctypes.c_double, dp, dp]
        L.orc_default_settings.argtypes = [ctypes.POINTER(OrcSettings)]
        vp = ctypes.POINTER(OrcVariant)
        L.orc_energy_ex.restype = ctypes.c_double
        L.orc_energy_ex.argtypes = [dp, ctypes.c_int, lp, ctypes.c_int, vp]
        L.orc_constraints_ex.restype = ctypes.c_int
        L.orc_constraints_ex.argtypes = [dp, ctypes.c_int, lp, dp, dp, vp, dp]
        L.orc_constraint_jacobian_ex.restype = ctypes.c_int
        L.orc_constraint_jacobian_ex.argtypes = [dp, ctypes.c_int, lp, ctypes.c_int, vp, dp]
        L.orc_objective_ex.restype = ctypes.c_double
        L.orc_objective_ex.argtypes = [dp, ctypes.c_int, lp, dp, dp, dp, ctypes.c_double, vp, dp, dp]
        L.orc_tree_dot.restype = ctypes.c_double
        L.orc_tree_dot.argtypes = [dp, dp, ctypes.c_int]
        L.orc_lbfgs_direction.argtypes = [dp, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int, dp]
        L.orc_lbfgs_memory.restype = ctypes.c_int
        L.orc_lbfgs_memory.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_solve_pixel.restype = ctypes.c_int
        L.orc_solve_pixel.argtypes = [dp, ctypes.c_int, lp, ctypes.c_int, dp, dp, dp,
                                      ctypes.POINTER(OrcSettings), dp, ctypes.POINTER(OrcCounters)]
        L.orc_solve_pixel_warm.restype = ctypes.c_int
        L.orc_solve_pixel_warm.argtypes = [dp, ctypes.c_int, lp, ctypes.c_int, dp, dp, ctypes.POINTER(OrcSettings),
                                           dp, ctypes.POINTER(OrcCounters)]
        L.orc_minimize.restype = ctypes.c_int
        L.orc_minimize.argtypes = [dp, dp, dp, ctypes.c_int, ORC_FN, ctypes.c_void_p,
                                   ctypes.POINTER(OrcSettings), dp, ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]
        L.orc_box.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.orc_guided_filter.argtypes = [dp, dp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int, dp]
        L.orc_default_radius.restype = ctypes.c_int
        L.orc_default_radius.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_unmix.restype = ctypes.c_int
        L.orc_unmix.argtypes = [dp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, lp, ctypes.c_int,
                                ctypes.c_int, ctypes.POINTER(OrcSettings), ctypes.c_int, dp,
                                ctypes.POINTER(OrcCounters)]
        L.orc_refine_targets.restype = ctypes.c_int
        L.orc_refine_targets.argtypes = [dp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, lp,
                                         ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, dp, dp, dp]
        L.orc_refine.restype = ctypes.c_int
        L.orc_refine.argtypes = [dp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, lp, ctypes.c_int,
                                 dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, dp, ctypes.POINTER(OrcSettings), ctypes.c_int, dp,
                                 ctypes.POINTER(OrcCounters)]
        L.orc_guided_filter_exact.argtypes = [dp, dp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.orc_qexp.restype = ctypes.c_int
        L.orc_qexp.argtypes = [ctypes.c_double]
        L.orc_qscale.restype = ctypes.c_int
        L.orc_qscale.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_qprefix.restype = ctypes.c_int
        L.orc_qprefix.argtypes = [ctypes.c_int]
        L.orc_coeff_bounds.restype = None
        L.orc_coeff_bounds.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.orc_recomposite.argtypes = [dp, ctypes.c_int, ctypes.c_int, lp, ctypes.c_int, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Layer:
    """A layer: Porter-Duff factors, blend mode id, mean, row-major precision."""
    mode: int
    mean: tuple
    precision: tuple
    comp: tuple = (1, 1, 1)


def layer_array(layers):
    arr = (OrcLayer * len(layers))()
    for i, l in enumerate(layers):
        arr[i].comp_x, arr[i].comp_y, arr[i].comp_z = (int(v) for v in l.comp)
        arr[i].blend_mode = int(l.mode)
        for c in range(3):
            arr[i].mean[c] = float(l.mean[c])
        P = np.asarray(l.precision, dtype=np.float64).reshape(9)
        for k in range(9):
            arr[i].precision[k] = float(P[k])
    return arr


def default_settings(**overrides) -> OrcSettings:
    s = OrcSettings()
    lib().orc_default_settings(ctypes.byref(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def variant(gray_mask=0, sparsity=0, sigma=10.0):
    v = OrcVariant()
    v.gray_mask, v.sparsity, v.sigma = int(gray_mask), int(sparsity), float(sigma)
    return v


def energy_ex(x, layers, minimum_alpha, v):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_energy_ex(_dp(x), len(layers), layer_array(layers), int(minimum_alpha), ctypes.byref(v))


def constraints_ex(x, layers, target, target_alphas, v):
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    C = np.zeros(3 + 4 * MAX_LAYERS)
    ta = None if target_alphas is None else np.ascontiguousarray(target_alphas, dtype=np.float64)
    nc = lib().orc_constraints_ex(_dp(x), len(layers), layer_array(layers), _dp(t), None if ta is None else _dp(ta),
                                  ctypes.byref(v), _dp(C))
    return C[:nc]


def constraint_jacobian_ex(x, layers, refine, v):
    L = len(layers)
    x = np.ascontiguousarray(x, dtype=np.float64)
    nc = (3 + L if refine else 4) + 3 * bin(v.gray_mask & ((1 << L) - 1)).count("1")
    J = np.zeros(4 * L * nc)
    lib().orc_constraint_jacobian_ex(_dp(x), L, layer_array(layers), int(bool(refine)), ctypes.byref(v), _dp(J))
    return J.reshape(4 * L, nc)


def objective_ex(x, layers, target, target_alphas, lam, rho, v):
    L = len(layers)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    ta = None if target_alphas is None else np.ascontiguousarray(target_alphas, dtype=np.float64)
    g = np.zeros(4 * L)
    C = np.zeros(3 + 4 * MAX_LAYERS)
    f = lib().orc_objective_ex(_dp(x), L, layer_array(layers), _dp(t), None if ta is None else _dp(ta), _dp(lam),
                               float(rho), ctypes.byref(v), _dp(g), _dp(C))
    return f, g


def blend(s, d, mode):
    return lib().orc_blend(float(s), float(d), int(mode))


def blend_grad_s(s, d, mode):
    return lib().orc_blend_grad_s(float(s), float(d), int(mode))


def blend_grad_d(s, d, mode):
    return lib().orc_blend_grad_d(float(s), float(d), int(mode))


def composite(x, layers):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_composite_stack(_dp(x), len(layers), layer_array(layers), _dp(out))
    return out


def energy(x, layers, minimum_alpha):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_energy(_dp(x), len(layers), layer_array(layers), int(minimum_alpha))


def constraints(x, layers, target, target_alphas=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    C = np.zeros(3 + 4 * MAX_LAYERS)
    ta = None if target_alphas is None else np.ascontiguousarray(target_alphas, dtype=np.float64)
    nc = lib().orc_constraints(_dp(x), len(layers), layer_array(layers), _dp(t),
                               None if ta is None else _dp(ta), _dp(C))
    return C[:nc]


def energy_gradient(x, layers, minimum_alpha):
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.zeros(4 * len(layers))
    lib().orc_energy_gradient(_dp(x), len(layers), layer_array(layers), int(minimum_alpha), _dp(g))
    return g


def constraint_jacobian(x, layers, refine):
    L = len(layers)
    x = np.ascontiguousarray(x, dtype=np.float64)
    nc = 3 + L if refine else 4
    J = np.zeros(4 * L * nc)
    lib().orc_constraint_jacobian(_dp(x), L, layer_array(layers), int(bool(refine)), _dp(J))
    return J.reshape(4 * L, nc)


def objective(x, layers, target, target_alphas, lam, rho):
    L = len(layers)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    ta = None if target_alphas is None else np.ascontiguousarray(target_alphas, dtype=np.float64)
    g = np.zeros(4 * L)
    C = np.zeros(3 + 4 * MAX_LAYERS)
    f = lib().orc_objective(_dp(x), L, layer_array(layers), _dp(t), None if ta is None else _dp(ta),
                            _dp(lam), float(rho), _dp(g), _dp(C))
    return f, g


def solve_pixel(target, layers, opaque=True, init_colors=None, target_alphas=None, bg=None,
                settings=None):
    L = len(layers)
    t = np.ascontiguousarray(target, dtype=np.float64)
    out = np.zeros(4 * L)
    cnt = OrcCounters()
    ic = None if init_colors is None else np.ascontiguousarray(init_colors, dtype=np.float64)
    ta = None if target_alphas is None else np.ascontiguousarray(target_alphas, dtype=np.float64)
    b = None if bg is None else np.ascontiguousarray(bg, dtype=np.float64)
    st = settings or default_settings()
    lib().orc_solve_pixel(_dp(t), L, layer_array(layers), int(bool(opaque)),
                          None if ic is None else _dp(ic), None if ta is None else _dp(ta),
                          None if b is None else _dp(b), ctypes.byref(st), _dp(out), ctypes.byref(cnt))
    return out, cnt


def solve_pixel_warm(target, layers, x_init, lambda_init=None, opaque=True, settings=None):
    L = len(layers)
    t = np.ascontiguousarray(target, dtype=np.float64)
    xi = np.ascontiguousarray(x_init, dtype=np.float64)
    li = None if lambda_init is None else np.ascontiguousarray(lambda_init, dtype=np.float64)
    out = np.zeros(4 * L)
    cnt = OrcCounters()
    st = settings or default_settings()
    capped = lib().orc_solve_pixel_warm(_dp(t), L, layer_array(layers), int(bool(opaque)), _dp(xi),
                                        None if li is None else _dp(li), ctypes.byref(st), _dp(out),
                                        ctypes.byref(cnt))
    return out, cnt, bool(capped)


def tree_dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().orc_tree_dot(_dp(a), _dp(b), a.size)


def lbfgs_direction(S, Y, gamma, q):
    """H q from the pairs (rows of S, Y, oldest first) by the minimiser's two-loop recursion."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    d = np.zeros_like(q)
    lib().orc_lbfgs_direction(_dp(S), _dp(Y), S.shape[0], float(gamma), _dp(q), q.size, _dp(d))
    return d


def lbfgs_memory(n, max_evals=1000, memory=0):
    return lib().orc_lbfgs_memory(int(n), int(max_evals), int(memory))


def minimize(fn, x0, lb, ub, settings=None):
    """Run the inner minimiser on a Python objective fn(x) -> (f, grad)."""
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    lb = np.ascontiguousarray(lb, dtype=np.float64)
    ub = np.ascontiguousarray(ub, dtype=np.float64)
    n = x.size
    trace = []

    def cb(xp, gp, _):
        xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
        f, g = fn(xv)
        for i in range(n):
            gp[i] = float(g[i])
        trace.append((xv, float(f)))
        return float(f)

    cfn = ORC_FN(cb)
    f = ctypes.c_double()
    ev = ctypes.c_int32()
    it = ctypes.c_int32()
    st = settings or default_settings()
    lib().orc_minimize(_dp(x), _dp(lb), _dp(ub), n, cfn, None, ctypes.byref(st), ctypes.byref(f),
                       ctypes.byref(ev), ctypes.byref(it), None)
    return x, f.value, ev.value, it.value, trace


def box(img, r, brute):
    img = np.ascontiguousarray(img, dtype=np.float64)
    h, w = img.shape
    out = np.zeros_like(img)
    lib().orc_box(_dp(img), w, h, int(r), int(bool(brute)), _dp(out))
    return out


def guided_filter(p, rgb, r, eps, brute=False):
    p = np.ascontiguousarray(p, dtype=np.float64)
    rgb = np.ascontiguousarray(rgb, dtype=np.float64)
    h, w = p.shape
    q = np.zeros_like(p)
    lib().orc_guided_filter(_dp(p), _dp(rgb), 0, w, h, int(r), float(eps), int(bool(brute)), _dp(q))
    return q


def guided_filter_exact(p, rgb, r, eps, imax=0.0, pmax=0.0):
    """The device's exact-sum guided filter (bounds <= 0: from the inputs)."""
    p = np.ascontiguousarray(p, dtype=np.float64)
    rgb = np.ascontiguousarray(rgb, dtype=np.float64)
    h, w = p.shape
    q = np.zeros_like(p)
    lib().orc_guided_filter_exact(_dp(p), _dp(rgb), 0, w, h, int(r), float(eps), float(imax), float(pmax), _dp(q))
    return q


FILTERS = {"sliding": 0, "brute": 1, "exact": 2}


def _filter_mode(filter, brute):
    return FILTERS["brute"] if brute else FILTERS[filter]


def default_radius(w, h):
    return lib().orc_default_radius(int(w), int(h))


def unmix(rgb, layers, opaque=True, settings=None, threads=0):
    rgb = np.ascontiguousarray(rgb, dtype=np.float64)
    h, w, _ = rgb.shape
    L = len(layers)
    out = np.zeros((L, h, w, 4))
    cnt = OrcCounters()
    st = settings or default_settings()
    rc = lib().orc_unmix(_dp(rgb), 0, w, h, layer_array(layers), L, int(bool(opaque)),
                         ctypes.byref(st), int(threads), _dp(out), ctypes.byref(cnt))
    assert rc == 0
    return out, cnt


def refine_targets(rgb, layers, first, opaque=True, smooth=False, radius=-1, eps=1e-4, brute=False, filter="exact",
                   bounds=None):
    """filter: 'exact' (the device's exact-sum filter), 'sliding' or 'brute'
    (FP64 sums); bounds: [max|I|, max|p| of each filtered plane] or None."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float64)
    first = np.ascontiguousarray(first, dtype=np.float64)
    h, w, _ = rgb.shape
    L = len(layers)
    targets = np.zeros((L, h, w))
    bg = np.zeros((3, h, w))
    b = None if bounds is None else np.ascontiguousarray(bounds, dtype=np.float64)
    lib().orc_refine_targets(_dp(rgb), 0, w, h, layer_array(layers), L, _dp(first), int(bool(opaque)),
                             int(bool(smooth)), int(radius), float(eps), _filter_mode(filter, brute),
                             None if b is None else _dp(b), _dp(targets), _dp(bg))
    return targets, (bg if smooth else None)


def filter_bounds(rgb, first, smooth=False):
    """[max|I|, max|alpha_l| for each layer, (max|c| of the background channels)]
    of an image and its packed first pass: the exact filter's bounds."""
    rgb = np.asarray(rgb, dtype=np.float64)
    first = np.asarray(first, dtype=np.float64)
    b = [float(np.abs(rgb).max())] + [float(np.abs(first[l, ..., 3]).max()) for l in range(first.shape[0])]
    if smooth:
        b += [float(np.abs(first[0, ..., c]).max()) for c in range(3)]
    return np.array(b)


def refine(rgb, layers, first, opaque=True, smooth=False, radius=-1, eps=1e-4, brute=False,
           settings=None, threads=0, filter="exact", bounds=None):
    rgb = np.ascontiguousarray(rgb, dtype=np.float64)
    first = np.ascontiguousarray(first, dtype=np.float64)
    h, w, _ = rgb.shape
    L = len(layers)
    out = np.zeros((L, h, w, 4))
    cnt = OrcCounters()
    st = settings or default_settings()
    b = None if bounds is None else np.ascontiguousarray(bounds, dtype=np.float64)
    rc = lib().orc_refine(_dp(rgb), 0, w, h, layer_array(layers), L, _dp(first), int(bool(opaque)),
                          int(bool(smooth)), int(radius), float(eps), _filter_mode(filter, brute),
                          None if b is None else _dp(b), ctypes.byref(st), int(threads), _dp(out),
                          ctypes.byref(cnt))
    assert rc == 0
    return out, cnt


def recomposite(packed, layers):
    packed = np.ascontiguousarray(packed, dtype=np.float64)
    L, h, w, _ = packed.shape
    out = np.zeros((h, w, 4))
    lib().orc_recomposite(_dp(packed), w, h, layer_array(layers), L, _dp(out))
    return out
