"""GPU checks of batched solves (unblending_hip_*_batch_device).

A batch runs every image of one layer count in one persistent launch, its
waves taking pixel chunks of all those images from a device queue; each
pixel's arithmetic is that of the single-image path, so the bar is
bit-identical outputs against per-image calls (and, through them, the
oracle parity of the single-image tests), for any chunk size.  A HIP graph
of the batch calls is replayed several times, with eager work in between,
and every replay is checked against the eager result.
"""
import numpy as np
import pytest

from conftest import layer_infos, oracle_layers

pytestmark = pytest.mark.gpu


def _images(count=6, scale=0.05, seed0=0):
    from unblending_amd import synth
    return [synth.config5_image(seed0 + j, scale) for j in range(count)]


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _batch(wls, firsts=None, alphas=None):
    from unblending_amd import device
    out = []
    for j, wl in enumerate(wls):
        out.append(device.BatchImage(_dev(wl.rgb), layer_infos(wl.layers),
                                     first=None if firsts is None else firsts[j],
                                     alpha_planes=None if alphas is None else alphas[j], radius=wl.radius))
    return out


def test_batch_unmix_matches_single_images():
    import torch
    import unblending_amd as U
    from unblending_amd import device
    wls = _images()
    print("case batch unmix: " + ", ".join(wl.name for wl in wls), flush=True)
    imgs = _batch(wls)
    for im in imgs:
        im.alpha_planes = torch.empty(im.out.shape[:3], dtype=torch.float64, device="cuda")
    st = U.HipStats()
    device.unmix_batch(imgs, stats=st)
    torch.cuda.synchronize()
    assert st.pixels == sum(wl.width * wl.height for wl in wls)
    for wl, im in zip(wls, imgs):
        single = device.unmix(im.rgb, im.layer_infos)
        assert torch.equal(single, im.out), wl.name
        assert torch.equal(im.alpha_planes, single[..., 3]), wl.name


@pytest.mark.parametrize("chunk", [64, 300, 4096])
def test_batch_chunk_size_does_not_change_results(chunk):
    import torch
    import unblending_amd as U
    from unblending_amd import device
    wls = _images(4, 0.06, seed0=10)
    print("case batch chunk %d" % chunk, flush=True)
    a = _batch(wls)
    device.unmix_batch(a)
    b = _batch(wls)
    device.unmix_batch(b, params=U.hip_params(chunk_size=chunk))
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x.out, y.out)


def test_batch_refine_matches_single_images():
    import torch
    import unblending_amd as U
    from unblending_amd import device
    wls = _images(5, 0.08, seed0=20)
    print("case batch refine: " + ", ".join("%s r=%d" % (wl.name, wl.radius) for wl in wls), flush=True)
    first = _batch(wls)
    device.unmix_batch(first)
    imgs = _batch(wls, firsts=[f.out for f in first])
    device.refine_batch(imgs)
    torch.cuda.synchronize()
    for wl, f, im in zip(wls, first, imgs):
        single = device.refine(im.rgb, f.out, im.layer_infos, params=U.hip_params(radius=wl.radius))
        assert torch.equal(single, im.out), wl.name


def test_batch_refine_many_images_over_the_filter_streams():
    """Nine images (more than the four filter streams, so that streams take
    several images in turn, each into its stream's coefficient region), with
    the smoothed background's three extra channels: every image equals its
    single-image refinement."""
    import torch
    import unblending_amd as U
    from unblending_amd import device
    wls = _images(9, 0.05, seed0=60)
    print("case batch refine over streams: " + ", ".join("%s r=%d" % (wl.name, wl.radius) for wl in wls), flush=True)
    first = _batch(wls)
    device.unmix_batch(first)
    imgs = _batch(wls, firsts=[f.out for f in first])
    device.refine_batch(imgs, smooth=True)
    torch.cuda.synchronize()
    for wl, f, im in zip(wls, first, imgs):
        single = device.refine(im.rgb, f.out, im.layer_infos, True, True, params=U.hip_params(radius=wl.radius))
        assert torch.equal(single, im.out), wl.name


def test_batch_matches_oracle(oracle):
    from unblending_amd import device
    wls = _images(3, 0.03, seed0=30)
    print("case batch vs oracle: " + ", ".join(wl.name for wl in wls), flush=True)
    imgs = _batch(wls)
    device.unmix_batch(imgs)
    for wl, im in zip(wls, imgs):
        ref, _ = oracle.unmix(wl.rgb, oracle_layers(wl.layers), opaque=True)
        d = np.abs(im.out.cpu().numpy() - ref)
        print("  %s max|diff| %.3e" % (wl.name, d.max()), flush=True)
        assert d.max() <= 1e-3


def test_batch_graph_replays_match_eager():
    import torch
    from unblending_amd import device
    wls = _images(4, 0.05, seed0=40)
    print("case batch graph", flush=True)
    # eager reference
    first = _batch(wls)
    device.unmix_batch(first)
    ref_first = [f.out.clone() for f in first]
    imgs = _batch(wls, firsts=[f.out for f in first])
    device.refine_batch(imgs)
    ref_out = [im.out.clone() for im in imgs]
    torch.cuda.synchronize()
    # capture both passes (workspaces allocated beforehand)
    ws1 = device.batch_workspace(first, False)
    ws2 = device.batch_workspace(imgs, True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        device.unmix_batch(first, ws=ws1)
        device.refine_batch(imgs, ws=ws2)
    other = _dev(np.random.default_rng(0).uniform(0, 1, (9, 11, 3)))
    for rep in range(3):
        for f, im in zip(first, imgs):
            f.out.fill_(-1.0)
            im.out.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for j in range(len(wls)):
            assert torch.equal(first[j].out, ref_first[j]), (rep, j)
            assert torch.equal(imgs[j].out, ref_out[j]), (rep, j)
        # eager work between replays
        device.unmix(other, layer_infos(wls[0].layers))
        torch.cuda.synchronize()


def test_warm_started_batch_matches_single_images(oracle):
    """Per-set warm start in a batch: each image with a model of its own layer
    set starts from that model's prediction; an image without one from the
    default initialisation.  Every image's first pass equals the
    single-image warm (or plain) call bit for bit, the fallback counts add
    up, and a warm image equals the oracle's warm solve from the same
    prediction."""
    import torch
    import unblending_amd as U
    from unblending_amd import device, mlp, synth
    wls = _images(3, 0.06, seed0=30)
    print("case warm batch: " + ", ".join(wl.name for wl in wls), flush=True)
    models = []
    for j, wl in enumerate(wls[:2]):
        tw = [synth.config5_image(30 + j, 0.06, image_seed=900 + 7 * j + t) for t in range(2)]
        images = torch.stack([_dev(t.rgb) for t in tw]).contiguous()
        targets = torch.stack([device.unmix(images[t], layer_infos(wl.layers)) for t in range(2)]).contiguous()
        m = mlp.WarmStartModel(len(wl.layers), seed=5 + j)
        m.fit(images, targets, steps=300, batch=1024, lr=3e-3, chunk=100)
        models.append(m)
    imgs = _batch(wls)
    imgs[0].model, imgs[1].model = models
    st = U.HipStats()
    device.unmix_batch(imgs, params=U.hip_params(warm_start=models[0]), stats=st)
    torch.cuda.synchronize()
    fallback = 0
    for j, (wl, im) in enumerate(zip(wls, imgs)):
        if j < 2:
            s1 = U.HipStats()
            single = device.unmix(im.rgb, im.layer_infos, params=U.hip_params(warm_start=models[j]), stats=s1)
            fallback += s1.fallback_pixels
        else:
            single = device.unmix(im.rgb, im.layer_infos)
        assert torch.equal(single, im.out), wl.name
    assert st.fallback_pixels == fallback
    # one warm image against the oracle's warm solve from the same prediction
    wl, im = wls[0], imgs[0]
    L = len(wl.layers)
    pred = models[0].predict(im.rgb).cpu().numpy()
    ol = oracle_layers(wl.layers)
    out = im.out.cpu().numpy()
    warm_settings = oracle.default_settings()
    warm_settings.max_count = 8
    for (y, x) in [(0, 0), (wl.height // 2, wl.width // 3), (wl.height - 1, wl.width - 1)]:
        x0 = np.concatenate([pred[:, y, x, 3], pred[:, y, x, :3].reshape(-1)])
        sol, _, cap = oracle.solve_pixel_warm(wl.rgb[y, x], ol, x0, settings=warm_settings)
        if cap:
            sol, _ = oracle.solve_pixel(wl.rgb[y, x], ol)
        assert np.array_equal(out[:, y, x, 3], sol[:L])
        assert np.array_equal(out[:, y, x, :3], sol[L:].reshape(L, 3))
