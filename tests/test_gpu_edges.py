"""GPU parity at the edges of image shapes and values.

Single pixels, single rows and columns, two-pixel-wide strips, constant and
saturated images, and guided-filter radii far beyond the image (the clamped
box then covers the whole image, and the workgroup strips hold a handful of
output columns).  These are where the solve's chunk queue, the filter's tile
and strip clamping and the default radius (60 min(w, h) / 1000, 0 below 17
pixels: the identity) meet their limits.  The bar is that of the other
parity tests: the oracle's outputs bit for bit (the filter: within 1e-6 of
the FP64 brute-force filter and equal to the oracle's exact-sum one).
"""
import numpy as np
import pytest

from conftest import layer_infos, oracle_layers

pytestmark = pytest.mark.gpu

FILTER_TOL = 1e-6


def _layers():
    from unblending_amd import synth
    return synth.config(2, (8, 8)).layers  # the 5-layer set


def _both_passes(oracle, rgb, memory, radii):
    import unblending_amd as U
    layers = _layers()
    ol = oracle_layers(layers)
    st = oracle.default_settings(memory=memory)
    first = U.compute_color_unmixing(rgb, layer_infos(layers), True, params=U.hip_params(lbfgs_memory=memory))
    ref_first, _ = oracle.unmix(rgb, ol, opaque=True, settings=st)
    print("  first pass max|diff| %.3e" % np.abs(first - ref_first).max(), flush=True)
    assert np.array_equal(first, ref_first)
    for r in radii:
        params = U.hip_params(radius=r, eps=1e-4, lbfgs_memory=memory)
        second = U.perform_matte_refinement(rgb, ref_first, layer_infos(layers), True, False, params=params)
        ref, _ = oracle.refine(rgb, ol, ref_first, opaque=True, radius=r, eps=1e-4, settings=st)
        print("  refinement r=%d max|diff| %.3e" % (r, np.abs(second - ref).max()), flush=True)
        assert np.array_equal(second, ref), r


@pytest.mark.parametrize("memory", [0, 1])
@pytest.mark.parametrize("h,w", [(1, 1), (1, 37), (29, 1), (2, 65), (3, 3)])
def test_thin_images_both_passes(oracle, h, w, memory):
    rng = np.random.default_rng(1000 * h + w)
    rgb = rng.uniform(0, 1, (h, w, 3))
    print("case thin image %dx%d memory %d" % (w, h, memory), flush=True)
    _both_passes(oracle, rgb, memory, radii=(-1, 1, 5))


@pytest.mark.parametrize("kind", ["black", "white", "grey", "one colour", "two colours"])
def test_constant_and_saturated_images(oracle, kind):
    h, w = 6, 9
    if kind == "black":
        rgb = np.zeros((h, w, 3))
    elif kind == "white":
        rgb = np.ones((h, w, 3))
    elif kind == "grey":
        rgb = np.full((h, w, 3), 0.5)
    elif kind == "one colour":
        rgb = np.broadcast_to(np.array([1.0, 0.0, 0.25]), (h, w, 3)).copy()
    else:
        rgb = np.zeros((h, w, 3))
        rgb[:, : w // 2] = (0.0, 1.0, 1.0)
        rgb[:, w // 2:] = (1.0, 1.0, 0.0)
    print("case %s image" % kind, flush=True)
    _both_passes(oracle, rgb, 0, radii=(2, 20))


@pytest.mark.parametrize("w,h,r", [(1, 40, 3), (40, 1, 3), (1, 1, 2), (2, 2, 96), (3, 200, 100), (200, 3, 100),
                                   (2, 2, 480), (70, 5, 130)])
def test_filter_thin_images(oracle, w, h, r):
    import torch
    from unblending_amd import device
    eps = 1e-4
    rng = np.random.default_rng(w * 1000 + h + r)
    rgb = rng.uniform(0, 1, (h, w, 3))
    planes = [rng.uniform(0, 1, (h, w)) for _ in range(4)]
    print("case filter %dx%d r=%d" % (w, h, r), flush=True)
    out = device.guided_filter(torch.from_numpy(rgb).cuda(), [torch.from_numpy(p).cuda() for p in planes], r, eps)
    torch.cuda.synchronize()
    for p, q in zip(planes, out):
        qn = q.cpu().numpy()
        assert np.abs(qn - oracle.guided_filter(p, rgb, r, eps, brute=True)).max() <= FILTER_TOL
        assert np.array_equal(qn, oracle.guided_filter_exact(p, rgb, r, eps))
