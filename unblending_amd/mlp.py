"""Warm-start model for the first pass (no counterpart in the reference).

The reference starts every pixel's solve at alpha 0.5 and the layer's mean
colour (unblending/src/unblending.cpp:108-127).  This module trains a small
network on the solver's own first-pass results for a layer set and lets the
first pass start each pixel from the network's prediction instead.  Pixels
whose warm-started solve stops at the outer-iteration cap without meeting the
tolerances are solved again from the reference's starting point, so every
pixel's result is a point the plain solver accepts.

The network (libunblending.so, include/unblending_hip.h): a pixel's 3x3 RGB
neighbourhood -> three GELU layers of 128 -> 4L sigmoid outputs (R, G, B,
alpha per layer), BF16 matrix-core products with FP32 accumulation, Adam on
FP32 master weights.  Torch supplies the device buffers and the stream.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib, raise_for_status

INPUTS, IN_PAD, HIDDEN, OUT_PAD = 27, 32, 128, 32
SHAPES = [(IN_PAD, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, OUT_PAD)]
MAX_LAYERS = OUT_PAD // 4


def _torch():
    import torch
    return torch


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def param_count() -> int:
    return int(_lib.unblending_hip_mlp_param_count())


def initial_params(layers: int, seed: int = 0) -> np.ndarray:
    """He-uniform hidden layers, Glorot-uniform output layer, zero biases;
    the padding (input rows 27..31, output columns >= 4 * layers) is zero
    and receives zero gradient, so it stays zero."""
    outputs = 4 * layers
    rng = np.random.default_rng(seed)
    p = np.zeros(param_count(), dtype=np.float32)
    off = 0
    for l, (fi, fo) in enumerate(SHAPES):
        if l < 3:
            lim = math.sqrt(6.0 / (INPUTS if l == 0 else fi))
        else:
            lim = math.sqrt(6.0 / (fi + outputs))
        w = rng.uniform(-lim, lim, size=(fi, fo)).astype(np.float32)
        if l == 0:
            w[INPUTS:, :] = 0.0
        if l == 3:
            w[:, outputs:] = 0.0
        p[off:off + fi * fo] = w.reshape(-1)
        off += fi * fo + fo
    assert off == p.size
    return p


class WarmStartModel:
    """Device-resident network for `layers` layers (at most 8)."""

    def __init__(self, layers: int, device=None, seed: int = 0, params: np.ndarray | None = None):
        torch = _torch()
        if not 1 <= layers <= MAX_LAYERS:
            raise ValueError("the warm-start model predicts 1 to %d layers" % MAX_LAYERS)
        self.layers = int(layers)
        self.device = torch.device(device or "cuda")
        size = int(_lib.unblending_hip_mlp_model_size())
        self.buffer = torch.empty((size + 7) // 8, dtype=torch.float64, device=self.device)
        self.steps_done = 0
        self.set_params(initial_params(layers, seed) if params is None else params)

    def data_ptr(self) -> int:
        return self.buffer.data_ptr()

    def set_params(self, params: np.ndarray, stream=None):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != param_count():
            raise ValueError("expected %d parameters" % param_count())
        raise_for_status(_lib.unblending_hip_mlp_set_params(
            ctypes.c_void_p(self.data_ptr()), p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _stream_ptr(stream)))
        self.steps_done = 0

    def params(self, stream=None) -> np.ndarray:
        p = np.empty(param_count(), dtype=np.float32)
        raise_for_status(_lib.unblending_hip_mlp_get_params(
            ctypes.c_void_p(self.data_ptr()), p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _stream_ptr(stream)))
        return p

    def save(self, path: str):
        np.savez(path, params=self.params(), layers=np.int64(self.layers))

    @classmethod
    def load(cls, path: str, device=None) -> "WarmStartModel":
        with np.load(path, allow_pickle=False) as f:
            return cls(int(f["layers"]), device=device, params=f["params"])

    def train(self, images, targets, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
              fixed_samples=None, update: bool = True, gradients: bool = False, workspace=None, stream=None):
        """Adam steps on pixels drawn from images (count, h, w, 3) with their
        first passes targets (count, L, h, w, 4), both float64 device tensors.
        Returns the per-step losses (and the last step's gradient when asked)."""
        torch = _torch()
        count, h, w, _ = images.shape
        if tuple(targets.shape) != (count, self.layers, h, w, 4):
            raise ValueError("targets must have shape (count, layers, h, w, 4)")
        for t in (images, targets):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("images and targets must be contiguous float64 device tensors")
        if fixed_samples is not None and (fixed_samples.dtype != torch.int64 or fixed_samples.numel() != batch):
            raise ValueError("fixed_samples must hold batch int64 indices")
        ws_size = int(_lib.unblending_hip_mlp_train_workspace_size(int(batch)))
        if ws_size == 0:
            raise ValueError("batch must be a positive multiple of 64")
        if workspace is None:
            workspace = torch.empty((ws_size + 7) // 8, dtype=torch.float64, device=self.device)
        losses = np.zeros(max(steps, 1), dtype=np.float32)
        grad = np.zeros(param_count(), dtype=np.float32) if gradients else None
        st = _lib.unblending_hip_mlp_train(
            ctypes.c_void_p(self.data_ptr()), ctypes.c_void_p(images.data_ptr()), count, w, h,
            ctypes.c_void_p(targets.data_ptr()), self.layers, int(steps), self.steps_done, int(batch), float(lr),
            int(seed) & ((1 << 64) - 1),
            ctypes.c_void_p(fixed_samples.data_ptr()) if fixed_samples is not None else None, int(bool(update)),
            ctypes.c_void_p(workspace.data_ptr()), workspace.numel() * 8, _stream_ptr(stream),
            losses.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            grad.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if grad is not None else None)
        raise_for_status(st)
        if update:
            self.steps_done += int(steps)
        losses = losses[:steps]
        return (losses, grad) if gradients else losses

    def fit(self, images, targets, steps: int = 6000, batch: int = 4096, lr: float = 2e-3, final_lr: float = 5e-5,
            chunk: int = 500, seed: int = 0, stream=None):
        """Adam with a cosine decay of the learning rate from lr to final_lr."""
        torch = _torch()
        ws = torch.empty((int(_lib.unblending_hip_mlp_train_workspace_size(batch)) + 7) // 8, dtype=torch.float64,
                         device=self.device)
        out = []
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            frac = (done + 0.5 * k) / steps
            rate = final_lr + 0.5 * (lr - final_lr) * (1.0 + math.cos(math.pi * frac))
            out.append(self.train(images, targets, k, batch, rate, seed, workspace=ws, stream=stream))
            done += k
        return np.concatenate(out) if out else np.zeros(0, dtype=np.float32)

    def predict(self, rgb, out=None, workspace=None, stream=None):
        """Predicted packed first pass (L, h, w, 4) of one image (h, w, 3)."""
        torch = _torch()
        h, w, _ = rgb.shape
        if rgb.dtype != torch.float64 or not rgb.is_cuda or not rgb.is_contiguous():
            raise ValueError("rgb must be a contiguous float64 device tensor")
        if out is None:
            out = torch.empty((self.layers, h, w, 4), dtype=torch.float64, device=rgb.device)
        if workspace is None:
            n = (int(_lib.unblending_hip_mlp_predict_workspace_size()) + 7) // 8
            workspace = torch.empty(n, dtype=torch.float64, device=rgb.device)
        raise_for_status(_lib.unblending_hip_mlp_predict(
            ctypes.c_void_p(self.data_ptr()), ctypes.c_void_p(rgb.data_ptr()), w, h, self.layers,
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(workspace.data_ptr()), workspace.numel() * 8,
            _stream_ptr(stream)))
        return out


def fit_layer_set(images, layer_infos, steps: int = 6000, batch: int = 4096, lr: float = 2e-3, seed: int = 0,
                  opaque: bool = True, device=None) -> WarmStartModel:
    """Solves the first pass of each training image (count, h, w, 3) on the
    device and fits a model to the results."""
    torch = _torch()
    from . import device as dev
    count, h, w, _ = images.shape
    L = len(layer_infos)
    targets = torch.empty((count, L, h, w, 4), dtype=torch.float64, device=images.device)
    ws = dev.workspace(w, h, L, False, device=images.device)
    for j in range(count):
        dev.unmix(images[j], layer_infos, opaque=opaque, out=targets[j], ws=ws)
    model = WarmStartModel(L, device=device or images.device, seed=seed)
    model.fit(images, targets, steps=steps, batch=batch, lr=lr, seed=seed)
    return model
