"""The allocation-free guided filter (unblending_hip_guided_filter_device_noalloc)
against the workspace entry it varies (unblending_hip_guided_filter_device),
bit for bit, on two calls in a row -- the second on a smaller image, so that
it runs over what the first left in the shared device scratch -- and the
rejection of a call whose scratch would exceed the array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(rgb_t, planes_t, out_t, radius, eps):
    import unblending_amd as U
    import torch
    h, w, _ = rgb_t.shape
    C = len(planes_t)
    pin = (ctypes.c_void_p * C)(*[p.data_ptr() for p in planes_t])
    qout = (ctypes.c_void_p * C)(*[q.data_ptr() for q in out_t])
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return U.lib().unblending_hip_guided_filter_device_noalloc(
        ctypes.c_void_p(rgb_t.data_ptr()), w, h, pin, qout, C, int(radius), float(eps), 0, 0, stream, None)


def test_noalloc_filter_matches_workspace_entry(oracle):
    import torch
    import unblending_amd as U
    from unblending_amd import device
    for (w, h, r, eps, C) in [(700, 650, 16, 1e-4, 7), (130, 90, 5, 1e-3, 3)]:
        print("case noalloc filter %dx%d r=%d C=%d" % (w, h, r, C), flush=True)
        rng = np.random.default_rng(w + h + r)
        rgb = rng.uniform(0, 1, (h, w, 3))
        planes = [rng.uniform(0, 1, (h, w)) for _ in range(C)]
        rgb_t = torch.from_numpy(rgb).cuda()
        planes_t = [torch.from_numpy(p).cuda() for p in planes]
        want = device.guided_filter(rgb_t, planes_t, r, eps)
        got = [torch.full((h, w), float("nan"), dtype=torch.float64, device="cuda") for _ in range(C)]
        st = _noalloc(rgb_t, planes_t, got, r, eps)
        assert st == U.OK, U.lib().unblending_hip_last_error()
        torch.cuda.synchronize()
        for c in range(C):
            assert torch.equal(got[c], want[c])
        assert np.array_equal(got[0].cpu().numpy(), oracle.guided_filter_exact(planes[0], rgb, r, eps))


def test_noalloc_filter_rejects_oversize():
    import torch
    import unblending_amd as U
    w, h, C = 2048, 2048, 8   # 8 (4 C n + 32) bytes > 256 MiB
    rgb_t = torch.zeros((h, w, 3), dtype=torch.float64, device="cuda")
    planes_t = [torch.zeros((h, w), dtype=torch.float64, device="cuda") for _ in range(C)]
    out_t = [torch.zeros((h, w), dtype=torch.float64, device="cuda") for _ in range(C)]
    st = _noalloc(rgb_t, planes_t, out_t, 8, 1e-4)
    assert st == U.INVALID_ARGUMENT
    assert b"exceeds" in U.lib().unblending_hip_last_error()
