"""Device-resident calls over torch tensors on the current HIP device.

Torch provides device memory and the stream; the work is done by
libunblending.so's device entry points (include/unblending_hip.h), enqueued
on torch's current stream so torch events time them.  Shapes:
    rgb    (h, w, 3)     float64, contiguous
    first  (L, h, w, 4)  float64 packed first pass
    out    (L, h, w, 4)
    alpha  (L, h, w)     optional alpha planes
"""
from __future__ import annotations

import ctypes

from . import (BACKEND_HIP, HipImage, HipParams, HipStats, _composite_tables, _lib, hip_params, layer_array,
               options, raise_for_status)


def _torch():
    import torch
    return torch


def _check(t, shape, name):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError("%s must be a contiguous float64 device tensor" % name)
    if tuple(t.shape) != tuple(shape):
        raise ValueError("%s has shape %s, expected %s" % (name, tuple(t.shape), tuple(shape)))


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _empty(shape, device, stream):
    """A float64 tensor for work enqueued on `stream`: allocated on that
    stream, so torch's caching allocator does not hand its memory to another
    stream's allocation while the library's kernels still use it."""
    torch = _torch()
    if stream is None:
        return torch.empty(shape, dtype=torch.float64, device=device)
    with torch.cuda.stream(stream):
        return torch.empty(shape, dtype=torch.float64, device=device)


def workspace_size(w, h, L, refine, smooth=False, warm=False, params: HipParams | None = None) -> int:
    """Workspace bytes for the device calls with these params (None: defaults)."""
    del warm  # the warm start's prediction needs no workspace
    return int(_lib.unblending_hip_workspace_size_ex(int(w), int(h), int(L), int(bool(refine)), int(bool(smooth)),
                                                     ctypes.byref(params) if params is not None else None))


def workspace(w, h, L, refine, smooth=False, device=None, warm=False, params: HipParams | None = None, stream=None):
    n = max(1, (workspace_size(w, h, L, refine, smooth, warm, params) + 7) // 8)
    return _empty(n, device or "cuda", stream)


def unmix(rgb, layer_infos, opaque=True, out=None, alpha_planes=None, ws=None, params: HipParams | None = None,
          stream=None, stats: HipStats | None = None):
    torch = _torch()
    h, w, _ = rgb.shape
    L = len(layer_infos)
    _check(rgb, (h, w, 3), "rgb")
    if out is None:
        out = _empty((L, h, w, 4), rgb.device, stream)
    _check(out, (L, h, w, 4), "out")
    if alpha_planes is not None:
        _check(alpha_planes, (L, h, w), "alpha_planes")
    if ws is None:
        ws = workspace(w, h, L, False, device=rgb.device, params=params, stream=stream)
    opt = options(BACKEND_HIP, opaque, False)
    st = _lib.unblending_hip_unmix_device(
        ctypes.c_void_p(rgb.data_ptr()), w, h, layer_array(layer_infos), L, ctypes.byref(opt),
        ctypes.byref(params or hip_params()), ctypes.c_void_p(out.data_ptr()),
        ctypes.c_void_p(alpha_planes.data_ptr()) if alpha_planes is not None else None,
        ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream_ptr(stream),
        ctypes.byref(stats) if stats is not None else None)
    raise_for_status(st)
    return out


def refine(rgb, first, layer_infos, opaque=True, smooth=False, alpha_planes=None, out=None, ws=None,
           params: HipParams | None = None, stream=None, stats: HipStats | None = None):
    torch = _torch()
    h, w, _ = rgb.shape
    L = len(layer_infos)
    _check(rgb, (h, w, 3), "rgb")
    _check(first, (L, h, w, 4), "first")
    if alpha_planes is not None:
        _check(alpha_planes, (L, h, w), "alpha_planes")
    if out is None:
        out = _empty((L, h, w, 4), rgb.device, stream)
    _check(out, (L, h, w, 4), "out")
    if ws is None:
        ws = workspace(w, h, L, True, smooth, device=rgb.device, params=params, stream=stream)
    opt = options(BACKEND_HIP, opaque, smooth)
    st = _lib.unblending_hip_refine_device(
        ctypes.c_void_p(rgb.data_ptr()), w, h, layer_array(layer_infos), L,
        ctypes.c_void_p(first.data_ptr()),
        ctypes.c_void_p(alpha_planes.data_ptr()) if alpha_planes is not None else None,
        ctypes.byref(opt), ctypes.byref(params or hip_params()), ctypes.c_void_p(out.data_ptr()),
        ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream_ptr(stream),
        ctypes.byref(stats) if stats is not None else None)
    raise_for_status(st)
    return out


def guided_filter(rgb, planes, radius, eps, out=None, ws=None, stream=None, stats: HipStats | None = None,
                  block=0, rows=0):
    """q[c] = guided filter of planes[c] (each (h, w)) with guidance rgb.
    block / rows: launch shape (0: defaults).  With stats, the call records
    device times (and synchronises)."""
    torch = _torch()
    h, w, _ = rgb.shape
    _check(rgb, (h, w, 3), "rgb")
    C = len(planes)
    for i, p in enumerate(planes):
        _check(p, (h, w), "planes[%d]" % i)
    if out is None:
        out = [_empty((h, w), rgb.device, stream) for _ in range(C)]
    if ws is None:
        n = max(1, (int(_lib.unblending_hip_filter_workspace_size(w, h, C)) + 7) // 8)
        ws = _empty(n, rgb.device, stream)
    pin = (ctypes.c_void_p * C)(*[p.data_ptr() for p in planes])
    qout = (ctypes.c_void_p * C)(*[q.data_ptr() for q in out])
    st = _lib.unblending_hip_guided_filter_device(
        ctypes.c_void_p(rgb.data_ptr()), w, h, pin, qout, C, int(radius), float(eps), int(block), int(rows),
        ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream_ptr(stream),
        ctypes.byref(stats) if stats is not None else None)
    raise_for_status(st)
    return out


def composite(layers, layer_infos, out=None, stream=None):
    """Image-level recomposite of device layers (L, h, w, 4) with the layer
    infos' comp-ops and blend modes into (h, w, 4) RGBA, on the device."""
    torch = _torch()
    L, h, w, _ = layers.shape
    _check(layers, (L, h, w, 4), "layers")
    if out is None:
        out = _empty((h, w, 4), layers.device, stream)
    _check(out, (h, w, 4), "out")
    ops, md = _composite_tables([i.comp_op for i in layer_infos], [i.blend_mode for i in layer_infos])
    raise_for_status(_lib.unblending_hip_composite_device(ctypes.c_void_p(layers.data_ptr()), w, h, L, ops, md,
                                                          ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)))
    return out


class BatchImage:
    """One image of a batch: rgb (h, w, 3) on the device, its layer infos,
    its output (L, h, w, 4) (allocated when None) and, for refinement, its
    first pass (L, h, w, 4), optional alpha planes (L, h, w) and radius.
    model: a trained mlp.WarmStartModel of the image's layer set, used by a
    first pass whose params ask for the warm start."""

    def __init__(self, rgb, layer_infos, out=None, first=None, alpha_planes=None, radius=-1, model=None):
        torch = _torch()
        h, w, _ = rgb.shape
        L = len(layer_infos)
        _check(rgb, (h, w, 3), "rgb")
        if out is None:
            out = torch.empty((L, h, w, 4), dtype=torch.float64, device=rgb.device)
        _check(out, (L, h, w, 4), "out")
        if first is not None:
            _check(first, (L, h, w, 4), "first")
        if alpha_planes is not None:
            _check(alpha_planes, (L, h, w), "alpha_planes")
        self.rgb, self.layer_infos, self.out, self.first, self.alpha_planes = rgb, layer_infos, out, first, alpha_planes
        self.radius = int(radius)
        self.model = model
        self._layers = layer_array(layer_infos)

    def c_struct(self) -> HipImage:
        h, w, _ = self.rgb.shape
        im = HipImage()
        im.rgb = self.rgb.data_ptr()
        im.width, im.height = w, h
        im.layers = ctypes.cast(self._layers, ctypes.c_void_p)
        im.layer_count = len(self.layer_infos)
        im.first_pass = self.first.data_ptr() if self.first is not None else None
        im.alpha_planes = self.alpha_planes.data_ptr() if self.alpha_planes is not None else None
        im.output = self.out.data_ptr()
        im.filter_radius = self.radius
        im.mlp = self.model.data_ptr() if self.model is not None else None
        return im


def _batch_array(images):
    arr = (HipImage * len(images))()
    for i, b in enumerate(images):
        arr[i] = b.c_struct()
    return arr


def batch_workspace_size(images, refine, smooth=False, params: HipParams | None = None) -> int:
    return int(_lib.unblending_hip_batch_workspace_size_ex(_batch_array(images), len(images), int(bool(refine)),
                                                           int(bool(smooth)),
                                                           ctypes.byref(params) if params is not None else None))


def batch_workspace(images, refine, smooth=False, device=None, params: HipParams | None = None, stream=None):
    n = max(1, (batch_workspace_size(images, refine, smooth, params) + 7) // 8)
    return _empty(n, device or images[0].rgb.device, stream)


def unmix_batch(images, opaque=True, ws=None, params: HipParams | None = None, stream=None,
                stats: HipStats | None = None):
    """First pass of every BatchImage (outputs, and alpha planes when given)."""
    if ws is None:
        ws = batch_workspace(images, False, params=params, stream=stream)
    opt = options(BACKEND_HIP, opaque, False)
    st = _lib.unblending_hip_unmix_batch_device(
        _batch_array(images), len(images), ctypes.byref(opt), ctypes.byref(params or hip_params()),
        ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream_ptr(stream),
        ctypes.byref(stats) if stats is not None else None)
    raise_for_status(st)
    return [b.out for b in images]


def refine_batch(images, opaque=True, smooth=False, ws=None, params: HipParams | None = None, stream=None,
                 stats: HipStats | None = None):
    """Refinement of every BatchImage (each with its first pass)."""
    if ws is None:
        ws = batch_workspace(images, True, smooth, params=params, stream=stream)
    opt = options(BACKEND_HIP, opaque, smooth)
    st = _lib.unblending_hip_refine_batch_device(
        _batch_array(images), len(images), ctypes.byref(opt), ctypes.byref(params or hip_params()),
        ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream_ptr(stream),
        ctypes.byref(stats) if stats is not None else None)
    raise_for_status(st)
    return [b.out for b in images]
