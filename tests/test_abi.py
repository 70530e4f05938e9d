"""The C interface (include/unblending.h, include/unblending_hip.h) without a GPU.

The library loads, exports every function the headers declare and nothing
else, reports ABI version 1, and applies the reference's argument checks in
the reference's order with the reference's status codes
(unblending/sdk/unblending_c.cpp:50-92, unblending/src/unblending.cpp:508-534,
590-614).  Where no device is present a valid call reports the device error.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import unblending_amd as U
from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", "unblending.h"), os.path.join(ROOT, "include", "unblending_hip.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        names.update(re.findall(r"UNBLENDING_API\s+[\w\s\*]+?\b(\w+)\s*\(", text))
    return names


def test_exports_match_headers():
    names = declared_functions()
    assert {"unblending_abi_version", "unblending_unmix", "unblending_refine"} <= names
    out = subprocess.run(["nm", "-D", "--defined-only", U.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert names == exported, (names ^ exported)
    lib = ctypes.CDLL(U.LIB_PATH)
    for n in names:
        assert hasattr(lib, n)


def test_soname_and_version():
    out = subprocess.run(["readelf", "-d", U.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "libunblending.so.1" in out
    assert U.abi_version() == 1
    assert U.BACKEND_HIP == 3 and U.DEVICE_ERROR == 3


def layers(n=3, mode=0, comp=(1, 1, 1)):
    arr = (U.CLayer * n)()
    for i in range(n):
        arr[i].comp_x, arr[i].comp_y, arr[i].comp_z = comp
        arr[i].blend_mode = mode
        for c in range(3):
            arr[i].mean[c] = 0.5
        for k, v in enumerate((np.eye(3) * 100.0).ravel()):
            arr[i].precision[k] = v
    return arr


class Call:
    def __init__(self, w=4, h=3, L=3):
        self.rgb = np.full((h, w, 3), 0.5)
        self.view = U.RgbView(self.rgb.ctypes.data_as(U._dp), self.rgb.size, w, h, 0)
        self.layers = layers(L)
        self.L = L
        self.out = np.full(4 * w * h * L, -7.0)
        self.opt = U.options(U.BACKEND_HIP)

    def unmix(self, view=True, layers=True, L=None, opt=True, out=True, cap=None):
        return U.lib().unblending_unmix(ctypes.byref(self.view) if view else None,
                                        self.layers if layers else None, self.L if L is None else L,
                                        ctypes.byref(self.opt) if opt else None,
                                        self.out.ctypes.data_as(U._dp) if out else None,
                                        self.out.size if cap is None else cap)

    def refine(self, first=None, size=None, **kw):
        first = np.zeros(self.out.size) if first is None else first
        return U.lib().unblending_refine(ctypes.byref(self.view), self.layers, self.L,
                                         first.ctypes.data_as(U._dp), first.size if size is None else size,
                                         ctypes.byref(self.opt), self.out.ctypes.data_as(U._dp), self.out.size)


def test_null_and_dimension_checks():
    c = Call()
    assert c.unmix(view=False) == U.INVALID_ARGUMENT
    assert c.unmix(layers=False) == U.INVALID_ARGUMENT
    assert c.unmix(L=0) == U.INVALID_ARGUMENT
    assert c.unmix(L=2**31) == U.INVALID_ARGUMENT
    c.view.width = 0
    assert c.unmix() == U.INVALID_ARGUMENT
    c.view.width = 4
    c.view.height = -1
    assert c.unmix() == U.INVALID_ARGUMENT
    c.view.height = 3
    c.view.width, c.view.height = 2**16, 2**16  # width > INT_MAX / height
    assert c.unmix() == U.INVALID_ARGUMENT


def test_capacity_is_checked_before_null_output():
    c = Call()
    assert c.unmix(cap=c.out.size - 1) == U.BUFFER_TOO_SMALL
    assert c.unmix(out=False, cap=c.out.size - 1) == U.BUFFER_TOO_SMALL
    assert c.unmix(out=False) == U.INVALID_ARGUMENT


def test_option_checks():
    c = Call()
    c.opt.struct_size = 4
    assert c.unmix() == U.INVALID_ARGUMENT
    c.opt = U.options(U.BACKEND_HIP)
    c.opt.opaque_background = 2
    assert c.unmix() == U.INVALID_ARGUMENT
    c.opt = U.options(U.BACKEND_HIP)
    c.opt.smooth_background = -1
    assert c.unmix() == U.INVALID_ARGUMENT
    c.opt = U.options(U.BACKEND_HIP, target_concurrency=-1)
    assert c.unmix() == U.INVALID_ARGUMENT
    c.opt = U.options(7)
    assert c.unmix() == U.INVALID_ARGUMENT
    c.opt = U.options(U.BACKEND_CPU)
    assert c.unmix() == U.INVALID_ARGUMENT  # no CPU solver in this library
    c.opt = U.options(U.BACKEND_CUDA)
    assert c.unmix() == U.DEVICE_ERROR      # built without CUDA


def test_layer_checks():
    c = Call()
    c.layers[1].blend_mode = 13
    assert c.unmix() == U.INVALID_ARGUMENT
    c.layers = layers()
    c.layers[0].mean[1] = float("nan")
    assert c.unmix() == U.INVALID_ARGUMENT
    c.layers = layers()
    c.layers[2].precision[1] = 1.0  # asymmetric
    assert c.unmix() == U.INVALID_ARGUMENT
    c.layers = layers()
    for k, v in enumerate(np.diag([1.0, -1.0, 1.0]).ravel()):  # symmetric, not positive definite
        c.layers[1].precision[k] = v
    assert c.unmix() == U.INVALID_ARGUMENT
    c.layers = layers()
    c.layers[1].precision[4] = float("inf")
    assert c.unmix() == U.INVALID_ARGUMENT


def test_stride_and_buffer_checks():
    c = Call()
    c.view.row_stride = 11  # < 3 * width
    assert c.unmix() == U.INVALID_ARGUMENT
    c.view.row_stride = 0
    c.view.size = 30  # < (h - 1) * stride + 3w
    assert c.unmix() == U.INVALID_ARGUMENT
    c.view.size = c.rgb.size
    c.view.data = None
    assert c.unmix() == U.INVALID_ARGUMENT


def test_refine_checks():
    c = Call()
    assert c.refine(size=c.out.size - 4) == U.INVALID_ARGUMENT
    c.opt = U.options(U.BACKEND_HIP, opaque=False, smooth=True)
    assert c.refine() == U.INVALID_ARGUMENT  # smoothing needs an opaque background
    c.opt = U.options(U.BACKEND_HIP)
    c.layers[1].comp_x = 2  # mixed source-over and plus
    assert c.refine() == U.INVALID_ARGUMENT


def test_valid_call_without_device_reports_device_error_and_leaves_output():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present")
    c = Call()
    assert c.unmix() == U.DEVICE_ERROR
    assert np.all(c.out == -7.0)
    c.opt = U.options(U.BACKEND_AUTO)
    assert c.unmix() == U.DEVICE_ERROR
    assert c.refine() == U.DEVICE_ERROR
    assert np.all(c.out == -7.0)


def test_environment_backend(monkeypatch):
    for v, b in [("", U.BACKEND_AUTO), ("auto", U.BACKEND_AUTO), ("cpu", U.BACKEND_CPU), ("cuda", U.BACKEND_CUDA),
                 ("hip", U.BACKEND_HIP)]:
        monkeypatch.setenv("UNBLENDING_BACKEND", v)
        assert U.environment_backend() == b
    monkeypatch.setenv("UNBLENDING_BACKEND", "vulkan")
    with pytest.raises(ValueError):
        U.environment_backend()


def test_device_entry_points_validate_workspace():
    lib = U.lib()
    L = layers(3)
    assert lib.unblending_hip_workspace_size(4, 3, 3, 0, 0, 0) >= 32
    assert lib.unblending_hip_workspace_size(4, 3, 3, 1, 1, 0) > lib.unblending_hip_workspace_size(4, 3, 3, 1, 0, 0)
    # the warm start's prediction is one fused kernel: no extra workspace
    assert lib.unblending_hip_workspace_size(4, 3, 3, 0, 0, 1) == lib.unblending_hip_workspace_size(4, 3, 3, 0, 0, 0)
    assert lib.unblending_hip_workspace_size(0, 3, 3, 1, 0, 0) == 0
    opt = U.options(U.BACKEND_HIP)
    # null rgb, then a too-small workspace: rejected before any device work
    assert lib.unblending_hip_unmix_device(None, 4, 3, L, 3, ctypes.byref(opt), None, ctypes.c_void_p(16), None,
                                           ctypes.c_void_p(16), 4096, None, None) == U.INVALID_ARGUMENT
    assert lib.unblending_hip_unmix_device(ctypes.c_void_p(16), 4, 3, L, 3, ctypes.byref(opt), None,
                                           ctypes.c_void_p(16), None, ctypes.c_void_p(16), 1, None,
                                           None) == U.BUFFER_TOO_SMALL


def test_warm_start_model_interface_checks():
    lib = U.lib()
    assert lib.unblending_hip_mlp_param_count() == 41376
    assert lib.unblending_hip_mlp_model_size() == 3 * 4 * 41376 + 2 * 2 * (4096 + 16384 + 16384 + 4096)
    assert lib.unblending_hip_mlp_train_workspace_size(0) == 0
    assert lib.unblending_hip_mlp_train_workspace_size(100) == 0
    assert lib.unblending_hip_mlp_train_workspace_size(128) < lib.unblending_hip_mlp_train_workspace_size(256)
    assert lib.unblending_hip_mlp_predict_workspace_size() == 0  # one fused kernel
    buf = ctypes.c_void_p(16)  # never dereferenced: every call below fails its argument checks first
    # more than 8 layers, bad batch, missing buffers
    assert lib.unblending_hip_mlp_predict(buf, buf, 4, 4, 9, buf, buf, 1 << 30, None) == U.INVALID_ARGUMENT
    assert lib.unblending_hip_mlp_predict(None, buf, 4, 4, 3, buf, buf, 1 << 30, None) == U.INVALID_ARGUMENT
    assert lib.unblending_hip_mlp_predict(buf, buf, 0, 4, 3, buf, None, 0, None) == U.INVALID_ARGUMENT
    assert lib.unblending_hip_mlp_train(buf, buf, 1, 4, 4, buf, 3, 1, 0, 100, 1e-3, 0, None, 1, buf, 1 << 30,
                                        None, None, None) == U.INVALID_ARGUMENT
    assert lib.unblending_hip_mlp_train(buf, buf, 1, 4, 4, buf, 3, 1, 0, 128, 1e-3, 0, None, 1, buf, 16,
                                        None, None, None) == U.BUFFER_TOO_SMALL
    assert lib.unblending_hip_mlp_set_params(None, None, None) == U.INVALID_ARGUMENT


def _batch_images(sizes, n_layers):
    arr = (U.HipImage * len(sizes))()
    keep = []
    for i, ((w, h), L) in enumerate(zip(sizes, n_layers)):
        lay = layers(L)
        keep.append(lay)
        arr[i].rgb = 16
        arr[i].width, arr[i].height = w, h
        arr[i].layers = ctypes.cast(lay, ctypes.c_void_p)
        arr[i].layer_count = L
        arr[i].first_pass = 16
        arr[i].output = 16
        arr[i].filter_radius = -1
    return arr, keep


def test_batch_entry_points_validate():
    lib = U.lib()
    imgs, keep = _batch_images([(4, 3), (5, 7), (2, 2)], [3, 5, 3])
    ws1 = lib.unblending_hip_batch_workspace_size(imgs, 3, 0, 0)
    ws2 = lib.unblending_hip_batch_workspace_size(imgs, 3, 1, 0)
    ws3 = lib.unblending_hip_batch_workspace_size(imgs, 3, 1, 1)
    assert 0 < ws1 < ws2 < ws3
    assert lib.unblending_hip_batch_workspace_size(None, 3, 0, 0) == 0
    assert lib.unblending_hip_batch_workspace_size(imgs, 0, 0, 0) == 0
    opt = U.options(U.BACKEND_HIP)
    buf = ctypes.c_void_p(16)  # never dereferenced: every call below fails its argument checks first
    unmix, refine = lib.unblending_hip_unmix_batch_device, lib.unblending_hip_refine_batch_device
    assert unmix(None, 3, ctypes.byref(opt), None, buf, ws1, None, None) == U.INVALID_ARGUMENT
    assert unmix(imgs, 0, ctypes.byref(opt), None, buf, ws1, None, None) == U.INVALID_ARGUMENT
    assert unmix(imgs, 3, ctypes.byref(opt), None, buf, ws1 - 1, None, None) == U.BUFFER_TOO_SMALL
    negative = U.hip_params(lbfgs_memory=-1)
    assert unmix(imgs, 3, ctypes.byref(opt), ctypes.byref(negative), buf, ws1, None, None) == U.INVALID_ARGUMENT
    simple = U.hip_params(solver=1)
    assert unmix(imgs, 3, ctypes.byref(opt), ctypes.byref(simple), buf, ws1, None, None) == U.INVALID_ARGUMENT
    # image checks in image order: a bad layer set, a missing first pass, too many layers
    imgs[1].layer_count = 0
    assert unmix(imgs, 3, ctypes.byref(opt), None, buf, ws1, None, None) == U.INVALID_ARGUMENT
    imgs[1].layer_count = 5
    keep[1][2].precision[1] = 7.0  # asymmetric precision
    assert unmix(imgs, 3, ctypes.byref(opt), None, buf, ws1, None, None) == U.INVALID_ARGUMENT
    keep[1][2].precision[1] = 0.0
    imgs[2].first_pass = None
    assert refine(imgs, 3, ctypes.byref(opt), None, buf, ws3, None, None) == U.INVALID_ARGUMENT
    imgs[2].first_pass = 16
    smooth_transparent = U.options(U.BACKEND_HIP, opaque=False, smooth=True)
    assert refine(imgs, 3, ctypes.byref(smooth_transparent), None, buf, ws3, None, None) == U.INVALID_ARGUMENT
    nine, keep9 = _batch_images([(4, 3)], [9])
    assert unmix(nine, 1, ctypes.byref(opt), None, buf, 1 << 40, None, None) == U.DEVICE_ERROR
    assert "1 to 8 layers" in U.last_error()


def test_composite_argument_checks_before_any_device_work():
    """unblending_hip_composite: argument errors and the output capacity are
    checked on the host, before the device is touched."""
    lib = U.lib()
    L, h, w = 2, 3, 4
    packed = np.zeros((L, h, w, 4))
    out = np.zeros((h, w, 4))
    dp = ctypes.POINTER(ctypes.c_double)
    ops = (ctypes.c_int32 * 6)(1, 1, 1, 1, 1, 1)

    def call(modes, size=packed.size, cap=out.size, n=L, width=w, p=packed):
        md = (ctypes.c_int32 * len(modes))(*modes)
        return lib.unblending_hip_composite(p.ctypes.data_as(dp) if p is not None else None, size, width, h, n,
                                            ops, md, out.ctypes.data_as(dp), cap)

    assert call([0, 13]) == U.INVALID_ARGUMENT          # unknown blend mode
    assert call([0, -1]) == U.INVALID_ARGUMENT
    assert call([0, 1], size=packed.size - 1) == U.INVALID_ARGUMENT  # not the packed size
    assert call([0, 1], cap=out.size - 1) == U.BUFFER_TOO_SMALL
    assert call([0, 1], width=0) == U.INVALID_ARGUMENT
    assert call([0, 1], p=None) == U.INVALID_ARGUMENT
    assert call([0] * 17, n=17, size=17 * h * w * 4) == U.INVALID_ARGUMENT  # more than 16 layers
