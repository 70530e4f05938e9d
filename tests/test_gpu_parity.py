"""GPU parity: the HIP passes against the oracle on the same seeded inputs.

The device solve runs the same algorithm in the same FP64 operation order as
the oracle, and the guided filter takes its box means from exact integer sums
(the oracle's exact-sum filter), so both passes agree with the oracle to the
last bit; that is asserted, together with the tolerances of the path:
per-pixel alphas and colours within 1e-3 max abs, recomposite RMS difference
at most 1e-4, guided filter within 1e-6 of the FP64 brute-force filter.
"""
import numpy as np
import pytest

from conftest import layer_infos, oracle_layers

pytestmark = pytest.mark.gpu

ALPHA_COLOR_TOL = 1e-3
RECOMPOSITE_RMS_TOL = 1e-4
FILTER_TOL = 1e-6


def _report(name, gpu, ref):
    d = np.abs(gpu - ref)
    exact = float(np.mean(np.all(d.reshape(d.shape[0], -1, 4) == 0, axis=(0, 2))))
    print("%s: max|diff| %.3e  bit-identical pixels %.4f" % (name, d.max(), exact), flush=True)
    return d


def _workload(k, size):
    from unblending_amd import synth
    return synth.config(k, size)


# (solver kernel, L-BFGS memory): the persistent kernel with the default
# history (NLopt's default vector storage: every pair of an inner solve), with
# a 3-pair ring that wraps, and with the one-pair register minimiser (the fast
# mode); the one-pixel-per-lane kernel (one pair)
MODES = [(0, 0), (0, 3), (0, 1), (1, 1)]


@pytest.mark.parametrize("solver,memory", MODES)
@pytest.mark.parametrize("k,size", [(1, (64, 48)), (2, (40, 40)), (3, (32, 36))])
def test_unmix_matches_oracle(oracle, k, size, solver, memory):
    import unblending_amd as U
    wl = _workload(k, size)
    print("case unmix c%d %s solver %d memory %d" % (k, size, solver, memory), flush=True)
    gpu = U.compute_color_unmixing(wl.rgb, layer_infos(wl.layers), True,
                                   params=U.hip_params(solver=solver, lbfgs_memory=memory))
    ref, _ = oracle.unmix(wl.rgb, oracle_layers(wl.layers), opaque=True,
                          settings=oracle.default_settings(memory=memory))
    d = _report("unmix c%d" % k, gpu, ref)
    assert d.max() <= ALPHA_COLOR_TOL
    assert d.max() == 0.0
    rg = oracle.recomposite(gpu, oracle_layers(wl.layers))
    rr = oracle.recomposite(ref, oracle_layers(wl.layers))
    assert np.sqrt(np.mean((rg - rr) ** 2)) <= RECOMPOSITE_RMS_TOL


@pytest.mark.parametrize("solver,memory", MODES)
@pytest.mark.parametrize("k,size", [(1, (64, 48)), (2, (40, 40)), (3, (32, 36))])
def test_refine_matches_oracle(oracle, k, size, solver, memory):
    import unblending_amd as U
    wl = _workload(k, size)
    print("case refine c%d %s solver %d memory %d" % (k, size, solver, memory), flush=True)
    ol = oracle_layers(wl.layers)
    st = oracle.default_settings(memory=memory)
    first, _ = oracle.unmix(wl.rgb, ol, opaque=True, settings=st)
    params = U.hip_params(radius=wl.radius, eps=wl.eps, solver=solver, lbfgs_memory=memory)
    gpu = U.perform_matte_refinement(wl.rgb, first, layer_infos(wl.layers), True, False, params=params)
    ref, _ = oracle.refine(wl.rgb, ol, first, opaque=True, radius=wl.radius, eps=wl.eps, settings=st)
    d = _report("refine c%d" % k, gpu, ref)
    assert d.max() <= ALPHA_COLOR_TOL
    assert d.max() == 0.0
    rg = oracle.recomposite(gpu, ol)
    rr = oracle.recomposite(ref, ol)
    assert np.sqrt(np.mean((rg - rr) ** 2)) <= RECOMPOSITE_RMS_TOL


@pytest.mark.parametrize("w,h,r,eps", [(37, 23, 3, 1e-4), (64, 64, 8, 1e-4), (20, 9, 12, 1e-3),
                                       (300, 40, 32, 1e-3), (5, 7, 0, 1e-4), (130, 70, 100, 1e-4)])
def test_guided_filter_matches_oracle(oracle, w, h, r, eps):
    import torch
    from unblending_amd import device
    rng = np.random.default_rng(w * 1000 + h + r)
    rgb = rng.uniform(0, 1, (h, w, 3))
    planes = [rng.uniform(0, 1, (h, w)) * (0.5 + 0.25 * c) for c in range(4)]  # two launch groups
    print("case filter %dx%d r=%d" % (w, h, r), flush=True)
    rgb_t = torch.from_numpy(rgb).cuda()
    out = device.guided_filter(rgb_t, [torch.from_numpy(p).cuda() for p in planes], r, eps)
    torch.cuda.synchronize()
    for p, q in zip(planes, out):
        qn = q.cpu().numpy()
        ref = oracle.guided_filter(p, rgb, r, eps, brute=True)
        err = np.abs(qn - ref).max()
        print("  max|q - q_brute| %.3e" % err, flush=True)
        assert err <= FILTER_TOL
        assert np.array_equal(qn, oracle.guided_filter_exact(p, rgb, r, eps))


@pytest.mark.parametrize("block,rows", [(256, 7), (128, 5), (512, 0), (1024, 33), (256, 200), (128, 0)])
def test_guided_filter_launch_shapes(oracle, block, rows):
    """Every launch shape the autotuner may pick computes the same filter."""
    import torch
    from unblending_amd import device
    w, h, r, eps = 150, 90, 9, 1e-4
    rng = np.random.default_rng(block + rows)
    rgb = rng.uniform(0, 1, (h, w, 3))
    planes = [rng.uniform(0, 1, (h, w)) for _ in range(2)]
    print("case filter shape block %d rows %d" % (block, rows), flush=True)
    out = device.guided_filter(torch.from_numpy(rgb).cuda(), [torch.from_numpy(p).cuda() for p in planes], r, eps,
                               block=block, rows=rows)
    torch.cuda.synchronize()
    for p, q in zip(planes, out):
        ref = oracle.guided_filter(p, rgb, r, eps, brute=True)
        assert np.abs(q.cpu().numpy() - ref).max() <= FILTER_TOL
        assert np.array_equal(q.cpu().numpy(), oracle.guided_filter_exact(p, rgb, r, eps))


@pytest.mark.parametrize("w,h,r,block", [(389, 41, 32, 256), (390, 37, 17, 256), (391, 35, 30, 128),
                                         (645, 19, 31, 256), (388, 33, 5, 128)])
def test_guided_filter_strip_edges(oracle, w, h, r, block):
    """Several strips a row, widths of every residue mod 4 and radii of both
    parities: lanes whose K columns are all output columns store them as
    16-byte pairs, the others (at each strip's halo and the image's right
    edge) column by column; both must give the exact filter."""
    import torch
    from unblending_amd import device
    eps = 1e-3
    rng = np.random.default_rng(w + 7 * h + 31 * r + block)
    rgb = rng.uniform(0, 1, (h, w, 3))
    planes = [rng.uniform(0, 1, (h, w)) * (0.6 + 0.1 * c) for c in range(3)]
    print("case filter strip edges %dx%d r=%d block %d" % (w, h, r, block), flush=True)
    out = device.guided_filter(torch.from_numpy(rgb).cuda(), [torch.from_numpy(p).cuda() for p in planes], r, eps,
                               block=block)
    torch.cuda.synchronize()
    for p, q in zip(planes, out):
        assert np.array_equal(q.cpu().numpy(), oracle.guided_filter_exact(p, rgb, r, eps))


@pytest.mark.parametrize("w,h,r,block", [(130, 70, 100, 1024), (130, 70, 100, 512), (600, 50, 245, 0)])
def test_guided_filter_wide_strips(oracle, w, h, r, block):
    """Radii above 96 run workgroup strips of 256, 512 or 1024 threads; five
    channels make launches of one channel (pass B at 1024 threads) and of
    three and two (the other passes)."""
    import torch
    from unblending_amd import device
    eps = 1e-4
    rng = np.random.default_rng(w + h + r + block)
    rgb = rng.uniform(0, 1, (h, w, 3))
    planes = [rng.uniform(0, 1, (h, w)) * (0.4 + 0.15 * c) for c in range(5)]
    print("case wide filter %dx%d r=%d block %d" % (w, h, r, block), flush=True)
    out = device.guided_filter(torch.from_numpy(rgb).cuda(), [torch.from_numpy(p).cuda() for p in planes], r, eps,
                               block=block)
    torch.cuda.synchronize()
    for p, q in zip(planes, out):
        qn = q.cpu().numpy()
        assert np.abs(qn - oracle.guided_filter(p, rgb, r, eps, brute=r <= 100)).max() <= FILTER_TOL
        assert np.array_equal(qn, oracle.guided_filter_exact(p, rgb, r, eps))


def test_refine_first_pass_may_alias_output():
    """The C ABI's refinement may take its first pass and its output in the
    same buffer (unblending/include/unblending.h:96-97): the result equals the
    call with separate buffers, bit for bit."""
    import ctypes
    import unblending_amd as U
    wl = _workload(2, (33, 29))
    infos = layer_infos(wl.layers)
    L = len(infos)
    print("case refine in place c2 33x29 L=%d" % L, flush=True)
    first = U.compute_color_unmixing(wl.rgb, infos, True)
    v, keep = U._view(wl.rgb)
    opt = U.options(U.BACKEND_HIP, True, False)
    dp = ctypes.POINTER(ctypes.c_double)
    separate = np.empty_like(first)
    st = U.lib().unblending_refine(ctypes.byref(v), U.layer_array(infos), L, first.ctypes.data_as(dp), first.size,
                                   ctypes.byref(opt), separate.ctypes.data_as(dp), separate.size)
    assert st == U.OK
    buf = first.copy()
    st = U.lib().unblending_refine(ctypes.byref(v), U.layer_array(infos), L, buf.ctypes.data_as(dp), buf.size,
                                   ctypes.byref(opt), buf.ctypes.data_as(dp), buf.size)
    assert st == U.OK
    assert np.array_equal(buf, separate)
    assert not np.array_equal(separate, first)


def test_host_calls_repeat_and_release_kept_buffers():
    """The host-pointer calls keep their device buffers between calls: calls
    of different sizes in turn (the kept buffers grown, then reused larger
    than needed), and calls after unblending_hip_release_buffers, all give
    the first results."""
    import unblending_amd as U
    work = [_workload(1, (40, 36)), _workload(2, (64, 60)), _workload(3, (20, 18))]
    print("case kept host buffers", flush=True)
    ref = []
    for wl in work:
        a = U.compute_color_unmixing(wl.rgb, layer_infos(wl.layers), True)
        ref.append((a, U.perform_matte_refinement(wl.rgb, a, layer_infos(wl.layers), True, False)))
    for rep in range(2):
        for wl, (a0, b0) in zip(work, ref):
            a = U.compute_color_unmixing(wl.rgb, layer_infos(wl.layers), True)
            assert np.array_equal(a, a0), (rep, wl.name)
            assert np.array_equal(U.perform_matte_refinement(wl.rgb, a, layer_infos(wl.layers), True, False), b0)
        U.release_device_buffers()
    U.release_device_buffers()  # with nothing kept


def test_concurrent_calls_from_two_threads():
    """The C ABI is re-entrant (each call has its own stream and buffers, and
    a call that finds the shared pinned staging buffers in use stages through
    its own): two host threads unmixing and refining different images at
    once get the results of the same calls made one after the other."""
    import threading
    import unblending_amd as U
    work = [_workload(1, (40, 36)), _workload(2, (28, 44))]
    print("case two threads c1 40x36 + c2 28x44", flush=True)
    alone = [U.compute_color_unmixing(wl.rgb, layer_infos(wl.layers), True) for wl in work]
    alone2 = [U.perform_matte_refinement(wl.rgb, a, layer_infos(wl.layers), True, False)
              for wl, a in zip(work, alone)]
    got = [None, None]
    got2 = [None, None]
    errors = []

    def run(i):
        try:
            for _ in range(3):
                got[i] = U.compute_color_unmixing(work[i].rgb, layer_infos(work[i].layers), True)
                got2[i] = U.perform_matte_refinement(work[i].rgb, got[i], layer_infos(work[i].layers), True, False)
        except Exception as e:  # reported below, after the join
            errors.append(e)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for i in range(2):
        assert np.array_equal(got[i], alone[i])
        assert np.array_equal(got2[i], alone2[i])
