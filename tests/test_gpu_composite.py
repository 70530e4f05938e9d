"""GPU: the image-level recomposite (unblending_hip_composite*, the device
counterpart of unblending/src/unblending.cpp:638-664) against the oracle's
recomposite of the same packed layers, bit for bit: every blend mode and both
comp-ops in one stack, host and device entries agreeing."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stack(rng, L, h, w):
    packed = rng.uniform(0, 1, (L, h, w, 4))
    packed[0, ..., 3] = 1.0
    packed[1, :3, :, 3] = 0.0  # transparent rows: the composite's A <= 1e-12 branch
    return packed


def test_composite_every_mode_matches_oracle(oracle):
    import unblending_amd as U
    rng = np.random.default_rng(11)
    modes = [0] + list(range(13))  # Normal background, then each of the 13 modes
    L, h, w = len(modes), 9, 31
    packed = _stack(rng, L, h, w)
    comps = [(1, 1, 1) if l % 3 else (2, 1, 1) for l in range(L)]
    print("case composite L=%d %dx%d" % (L, w, h), flush=True)
    infos = [U.LayerInfo(U.CompOp(*comps[l]), U.BlendMode(modes[l]),
                         U.GaussianColorModel((0.5, 0.5, 0.5), np.eye(3))) for l in range(L)]
    got = U.composite_layers(packed, infos)
    ol = [oracle.Layer(modes[l], (0.5, 0.5, 0.5), tuple(np.eye(3).ravel()), comps[l]) for l in range(L)]
    ref = oracle.recomposite(packed, ol)
    assert np.array_equal(got, ref)


def test_composite_device_entry_matches_host_entry():
    import torch
    import unblending_amd as U
    from unblending_amd import device
    rng = np.random.default_rng(12)
    L, h, w = 5, 40, 33
    packed = _stack(rng, L, h, w)
    infos = [U.LayerInfo(U.CompOp.SourceOver(), U.BlendMode(m), U.GaussianColorModel((0.5, 0.5, 0.5), np.eye(3)))
             for m in (0, 1, 2, 6, 3)]
    print("case composite device %dx%d" % (w, h), flush=True)
    host = U.composite_layers(packed, infos)
    dev = device.composite(torch.from_numpy(packed).cuda(), infos)
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy(), host)


def test_recomposite_of_the_passes_reconstructs_the_image(oracle):
    """The composite of both passes' layers is the input image within the
    solver's tolerance (the reconstruction the north-star RMS bar measures)."""
    import unblending_amd as U
    from unblending_amd import synth
    from conftest import layer_infos
    wl = synth.config(1, (48, 40))
    infos = layer_infos(wl.layers)
    print("case recomposite c1 48x40", flush=True)
    first = U.compute_color_unmixing(wl.rgb, infos, True)
    second = U.perform_matte_refinement(wl.rgb, first, infos, True, False,
                                        params=U.hip_params(radius=wl.radius, eps=wl.eps))
    for packed in (first, second):
        rgba = U.composite_layers(packed, infos)
        rms = float(np.sqrt(np.mean((rgba[..., :3] - wl.rgb) ** 2)))
        print("  reconstruction RMS %.3e" % rms, flush=True)
        assert rms < 5e-3
        assert np.abs(rgba[..., 3] - 1.0).max() < 1e-12  # opaque background
