"""GPU parity at every layer count the device compiles (1 to 8), both passes.

The search-direction service splits a pixel's 4L variables over four lanes,
eight components each: which lanes hold components, and which halves of a
lane's share exist, depend on L (L = 1 and 2 leave three lanes without any,
L = 3, 5 and 7 a lane with half a share).  Each case solves a small seeded
image whose layer set cycles through all thirteen blend modes with full SPD
colour models, with the default minimiser (NLopt's default history) and the
one-pair fast mode, and requires the oracle's outputs bit for bit.
"""
import numpy as np
import pytest

from conftest import layer_infos, oracle_layers

pytestmark = pytest.mark.gpu


def _workload(L, w=20, h=16, seed=0x5EED0800):
    from unblending_amd import synth
    rng = np.random.default_rng(seed + L)
    modes = [synth.NORMAL] + [(1 + (L + k) % 12) for k in range(L - 1)]
    layers = [synth.spd_model(rng, m, 0.1, 0.9, 0.002, 0.02) for m in modes]
    fields = [None] + [synth.blurred_noise(rng, h, w, 3.0) for _ in modes[1:]]
    rgb = synth._render(rng, layers, fields, 0.005, h, w)
    return synth.Workload("layers %d" % L, rgb, layers, 3, 1e-4)


@pytest.mark.parametrize("memory", [0, 1])
@pytest.mark.parametrize("L", list(range(1, 9)))
def test_both_passes_match_oracle(oracle, L, memory):
    import unblending_amd as U
    wl = _workload(L)
    print("case layers %d memory %d" % (L, memory), flush=True)
    ol = oracle_layers(wl.layers)
    st = oracle.default_settings(memory=memory)
    params = U.hip_params(radius=wl.radius, eps=wl.eps, lbfgs_memory=memory)
    first = U.compute_color_unmixing(wl.rgb, layer_infos(wl.layers), True, params=params)
    ref_first, _ = oracle.unmix(wl.rgb, ol, opaque=True, settings=st)
    print("  first pass max|diff| %.3e" % np.abs(first - ref_first).max(), flush=True)
    assert np.array_equal(first, ref_first)
    second = U.perform_matte_refinement(wl.rgb, ref_first, layer_infos(wl.layers), True, False, params=params)
    ref_second, _ = oracle.refine(wl.rgb, ol, ref_first, opaque=True, radius=wl.radius, eps=wl.eps, settings=st)
    print("  refinement max|diff| %.3e" % np.abs(second - ref_second).max(), flush=True)
    assert np.array_equal(second, ref_second)


@pytest.mark.parametrize("L", [2, 6, 8])
def test_non_opaque_first_pass_matches_oracle(oracle, L):
    """Without the opaque bottom layer the first alpha is a free variable."""
    import unblending_amd as U
    wl = _workload(L, seed=0x5EED0900)
    print("case non-opaque layers %d" % L, flush=True)
    first = U.compute_color_unmixing(wl.rgb, layer_infos(wl.layers), False)
    ref, _ = oracle.unmix(wl.rgb, oracle_layers(wl.layers), opaque=False)
    assert np.array_equal(first, ref)
