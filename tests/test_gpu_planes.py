"""The planar seam uhip_solve_planes (the counterpart of the reference's
gpu::solve_cuda, unblending/src/cuda/pixel_pipeline.cu:33-107, called from
unblending/src/gpu_dispatch.cpp:88-110 with rgb[3][n] and, for refinement,
initial[4L][n] -- the L alpha planes, then the colour planes at
(L + 3 l + c) n -- and repacked at gpu_dispatch.cpp:182-191): both passes
through the planar layout, against the oracle after repacking, with the
reference's default filter radius 60 min(w, h) / 1000 and eps 1e-4."""
import ctypes

import numpy as np
import pytest

from conftest import layer_infos, oracle_layers


def _c_layers(layers):
    import unblending_amd as U
    return U.layer_array(layer_infos(layers))


def to_planes(packed):
    """[L][h][w][RGBA] -> [4L][n]: alpha planes, then colour planes."""
    L = packed.shape[0]
    n = packed.shape[1] * packed.shape[2]
    flat = packed.reshape(L, n, 4)
    return np.ascontiguousarray(np.concatenate([flat[:, :, 3], flat[:, :, :3].transpose(0, 2, 1).reshape(3 * L, n)]))


def from_planes(planes, L, h, w):
    n = h * w
    out = np.empty((L, n, 4))
    out[:, :, 3] = planes[:L]
    out[:, :, :3] = planes[L:].reshape(L, 3, n).transpose(0, 2, 1)
    return out.reshape(L, h, w, 4)


def _solve_planes(wl, first_planes, opaque, smooth):
    import unblending_amd as U
    h, w, _ = wl.rgb.shape
    L = len(wl.layers)
    n = h * w
    rgb_planes = np.ascontiguousarray(wl.rgb.reshape(n, 3).T)
    out = np.full((4 * L, n), np.nan)
    dp = ctypes.POINTER(ctypes.c_double)
    st = U.lib().uhip_solve_planes(w, h, _c_layers(wl.layers), L, rgb_planes.ctypes.data_as(dp),
                                   first_planes.ctypes.data_as(dp) if first_planes is not None else None,
                                   int(opaque), int(smooth), out.ctypes.data_as(dp))
    return st, out


@pytest.mark.gpu
@pytest.mark.parametrize("k,size,smooth", [(1, (64, 48), False), (2, (40, 40), False), (2, (40, 40), True)])
def test_planar_seam_matches_oracle(oracle, k, size, smooth):
    import unblending_amd as U
    from unblending_amd import synth
    wl = synth.config(k, size)
    L = len(wl.layers)
    h, w, _ = wl.rgb.shape
    print("case planes c%d %s smooth %d" % (k, size, smooth), flush=True)
    ol = oracle_layers(wl.layers)
    st, p1 = _solve_planes(wl, None, True, False)
    assert st == U.OK, U.lib().unblending_hip_last_error()
    first = from_planes(p1, L, h, w)
    ref1, _ = oracle.unmix(wl.rgb, ol, opaque=True)
    assert np.array_equal(first, ref1)
    st, p2 = _solve_planes(wl, to_planes(first), True, smooth)
    assert st == U.OK, U.lib().unblending_hip_last_error()
    second = from_planes(p2, L, h, w)
    ref2, _ = oracle.refine(wl.rgb, ol, ref1, opaque=True, smooth=smooth, radius=-1, eps=1e-4)
    assert oracle.default_radius(w, h) == 60 * min(w, h) // 1000
    assert np.array_equal(second, ref2)
    # the same as the packed C ABI
    packed = U.perform_matte_refinement(wl.rgb, first, layer_infos(wl.layers), True, smooth)
    assert np.array_equal(packed, second)


@pytest.mark.gpu
def test_planar_seam_threaded_repack_matches_packed():
    """At 1024 x 640 the seam's repacks run on several host threads (ranges
    of at least 2^18 pixels); both passes through the planes equal the packed
    C interface bit for bit."""
    import unblending_amd as U
    from unblending_amd import synth
    wl = synth.config(1, (1024, 640))
    L = len(wl.layers)
    h, w, _ = wl.rgb.shape
    infos = layer_infos(wl.layers)
    print("case planes threaded repack %dx%d" % (w, h), flush=True)
    st, p1 = _solve_planes(wl, None, True, False)
    assert st == U.OK, U.lib().unblending_hip_last_error()
    first = from_planes(p1, L, h, w)
    assert np.array_equal(first, U.compute_color_unmixing(wl.rgb, infos, True))
    st, p2 = _solve_planes(wl, to_planes(first), True, False)
    assert st == U.OK, U.lib().unblending_hip_last_error()
    assert np.array_equal(from_planes(p2, L, h, w), U.perform_matte_refinement(wl.rgb, first, infos, True, False))


def test_planar_seam_rejects_bad_arguments_before_allocating():
    import unblending_amd as U
    from unblending_amd import synth
    wl = synth.config(1, (8, 6))
    L = len(wl.layers)
    dp = ctypes.POINTER(ctypes.c_double)
    rgb = np.zeros((3, 48))
    out = np.zeros((4 * L, 48))
    lib = U.lib()
    lays = _c_layers(wl.layers)
    # width * height beyond INT_MAX: rejected by the dimension check, nothing allocated
    assert lib.uhip_solve_planes(1 << 20, 1 << 20, lays, L, rgb.ctypes.data_as(dp), None, 1, 0,
                                 out.ctypes.data_as(dp)) == U.INVALID_ARGUMENT
    assert lib.uhip_solve_planes(0, 6, lays, L, rgb.ctypes.data_as(dp), None, 1, 0,
                                 out.ctypes.data_as(dp)) == U.INVALID_ARGUMENT
    assert lib.uhip_solve_planes(8, 6, lays, 0, rgb.ctypes.data_as(dp), None, 1, 0,
                                 out.ctypes.data_as(dp)) == U.INVALID_ARGUMENT
    assert lib.uhip_solve_planes(8, 6, lays, L, None, None, 1, 0,
                                 out.ctypes.data_as(dp)) == U.INVALID_ARGUMENT
    # a layer count that overflows the packed size
    assert lib.uhip_solve_planes(8, 6, lays, (1 << 62), rgb.ctypes.data_as(dp), None, 1, 0,
                                 out.ctypes.data_as(dp)) == U.INVALID_ARGUMENT
