"""GPU parity for the problem variants the passes accept besides the
default opaque source-over stack: a smoothed background in refinement (three
more filtered channels, background colour fixed; unblending.cpp:392-399,
174-208), all-Plus stacks (targets normalised to sum to one;
unblending.cpp:311-333), a non-opaque first pass (alpha_0 free), and the
reference's compile-time variants: layer 1 constrained to gray
(SECOND_LAYER_GRAY) and the first pass's sparsity term (SPARCITY).  Each
against the oracle on the same inputs, bit for bit."""
import dataclasses

import numpy as np
import pytest

from conftest import layer_infos, oracle_layers

pytestmark = pytest.mark.gpu


def _plus(layers):
    return [dataclasses.replace(l, comp=(2, 1, 1)) for l in layers]


@pytest.mark.parametrize("k,size", [(1, (40, 36)), (2, (36, 30))])
def test_refine_smooth_background_matches_oracle(oracle, k, size):
    import unblending_amd as U
    from unblending_amd import synth
    wl = synth.config(k, size)
    ol = oracle_layers(wl.layers)
    print("case refine smooth c%d %s" % (k, size), flush=True)
    first, _ = oracle.unmix(wl.rgb, ol, opaque=True)
    params = U.hip_params(radius=wl.radius, eps=wl.eps)
    gpu = U.perform_matte_refinement(wl.rgb, first, layer_infos(wl.layers), True, True, params=params)
    ref, _ = oracle.refine(wl.rgb, ol, first, opaque=True, smooth=True, radius=wl.radius, eps=wl.eps)
    print("  max|diff| %.3e" % np.abs(gpu - ref).max(), flush=True)
    assert np.array_equal(gpu, ref)
    # the background layer's colour is the smoothed one, held fixed
    assert np.all((gpu[0, ..., :3] >= 0.0) & (gpu[0, ..., :3] <= 1.0))


def test_smooth_without_opaque_is_rejected():
    import unblending_amd as U
    from unblending_amd import synth
    wl = synth.config(1, (16, 12))
    first = np.zeros((len(wl.layers), 12, 16, 4))
    with pytest.raises(ValueError):  # unblending.cpp:599-600
        U.perform_matte_refinement(wl.rgb, first, layer_infos(wl.layers), False, True)


def test_plus_stack_both_passes_match_oracle(oracle):
    import unblending_amd as U
    from unblending_amd import synth
    wl = synth.config(1, (40, 32))
    layers = _plus(wl.layers)
    ol = oracle_layers(layers)
    infos = layer_infos(layers)
    print("case plus stack", flush=True)
    gpu1 = U.compute_color_unmixing(wl.rgb, infos, False)
    ref1, _ = oracle.unmix(wl.rgb, ol, opaque=False)
    print("  unmix max|diff| %.3e" % np.abs(gpu1 - ref1).max(), flush=True)
    assert np.array_equal(gpu1, ref1)
    params = U.hip_params(radius=wl.radius, eps=wl.eps)
    gpu2 = U.perform_matte_refinement(wl.rgb, ref1, infos, False, False, params=params)
    ref2, _ = oracle.refine(wl.rgb, ol, ref1, opaque=False, radius=wl.radius, eps=wl.eps)
    print("  refine max|diff| %.3e" % np.abs(gpu2 - ref2).max(), flush=True)
    assert np.array_equal(gpu2, ref2)


def test_non_opaque_first_pass_matches_oracle(oracle):
    import unblending_amd as U
    from unblending_amd import synth
    wl = synth.config(2, (36, 28))
    ol = oracle_layers(wl.layers)
    print("case unmix non-opaque c2", flush=True)
    gpu = U.compute_color_unmixing(wl.rgb, layer_infos(wl.layers), False)
    ref, _ = oracle.unmix(wl.rgb, ol, opaque=False)
    print("  max|diff| %.3e" % np.abs(gpu - ref).max(), flush=True)
    assert np.array_equal(gpu, ref)


# The reference's compile-time problem variants: SECOND_LAYER_GRAY (layer 1's
# colour constrained to gray, both passes; unblending.cpp:21,168-172) and
# SPARCITY (the first pass's sparsity term; unblending.cpp:17,236-240).
VARIANTS = [(0b10, False), (0, True), (0b10, True)]


@pytest.mark.parametrize("gray,sparse", VARIANTS)
@pytest.mark.parametrize("k,size", [(1, (40, 32)), (3, (24, 20))])
def test_problem_variants_match_oracle(oracle, k, size, gray, sparse):
    import unblending_amd as U
    from unblending_amd import synth
    wl = synth.config(k, size)
    print("case variant c%d %s gray %d sparse %d" % (k, size, gray, sparse), flush=True)
    ol = oracle_layers(wl.layers)
    infos = layer_infos(wl.layers)
    st = oracle.default_settings(gray_mask=gray, sparsity=int(sparse))
    params = U.hip_params(radius=wl.radius, eps=wl.eps, gray_layers=gray, sparsity=sparse)
    first = U.compute_color_unmixing(wl.rgb, infos, True, params=params)
    ref1, _ = oracle.unmix(wl.rgb, ol, opaque=True, settings=st)
    assert np.abs(first - ref1).max() == 0.0
    second = U.perform_matte_refinement(wl.rgb, ref1, infos, True, False, params=params)
    ref2, _ = oracle.refine(wl.rgb, ol, ref1, opaque=True, radius=wl.radius, eps=wl.eps, settings=st)
    assert np.abs(second - ref2).max() == 0.0
    if gray:
        plain = U.compute_color_unmixing(wl.rgb, infos, True, params=U.hip_params())
        c1 = first[1, ..., :3]
        res = np.abs(np.sqrt(3.0) * c1 - np.linalg.norm(c1, axis=-1, keepdims=True)).max(-1)
        res0 = np.abs(np.sqrt(3.0) * plain[1, ..., :3] - np.linalg.norm(plain[1, ..., :3], axis=-1,
                                                                          keepdims=True)).max(-1)
        print("gray residual median %.2e (plain %.2e)" % (np.median(res), np.median(res0)), flush=True)
        assert np.median(res) < 5e-3 < np.median(res0)


def test_problem_variants_rejected_where_not_built():
    import unblending_amd as U
    from unblending_amd import synth
    wl = synth.config(1, (8, 6))
    infos = layer_infos(wl.layers)
    for p in (U.hip_params(gray_layers=0b100), U.hip_params(gray_layers=0b10, lbfgs_memory=1),
              U.hip_params(gray_layers=0b10, solver=1, lbfgs_memory=1)):
        with pytest.raises(ValueError):
            U.compute_color_unmixing(wl.rgb, infos, True, params=p)
