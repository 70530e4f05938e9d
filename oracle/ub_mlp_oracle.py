"""NumPy restatement of the warm-start network (unblending_amd/csrc/ub_mlp.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
network's forward pass, gradients, Adam update and batch sampling.  The
product never imports it.

The network has no counterpart in the reference (the reference starts every
pixel's solve from the same initial point, unblending/src/unblending.cpp:
108-127); it is an optional accelerator whose output only changes where the
solve starts.  What fixes each function here:

  gelu          tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)));
                known values and the exact-erf form within 1e-3
  forward       a plain fp64 evaluation of the stated architecture
  loss_grad     fp64 backpropagation, pinned by central finite differences
                of `loss` (a sign or index error in any layer fails them);
                its `plant` option computes the gradient with a named slip,
                for tests of how far such slips move it
  adam          Kingma & Ba, Algorithm 1; the first step is -lr * g/(|g|+eps)
                up to eps, a closed form
  splitmix64    Steele, Lea & Flood 2014; the first output of state 0 is the
                published 0xE220A8397B1DCDAF
  *_bf16        the same computations with operands rounded to bfloat16
                where the device rounds them (round to nearest even), for a
                tight comparison with the matrix-core path

Parameter vector (float, padded): W1 [32][128], b1 [128], W2 [128][128],
b2, W3 [128][128], b3, W4 [128][32], b4 [32]; weights [in][out].
"""
from __future__ import annotations

import numpy as np

INPUTS, IN_PAD, HIDDEN, OUT_PAD = 27, 32, 128, 32
SHAPES = [(IN_PAD, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, OUT_PAD)]
GELU_C0 = 0.7978845608028654
GELU_C1 = 0.044715


def _offsets():
    off, ow, ob = 0, [], []
    for (i, o) in SHAPES:
        ow.append(off)
        off += i * o
        ob.append(off)
        off += o
    return ow, ob, off


OFF_W, OFF_B, N_PARAMS = _offsets()


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    W = [p[OFF_W[l]:OFF_W[l] + i * o].reshape(i, o) for l, (i, o) in enumerate(SHAPES)]
    b = [p[OFF_B[l]:OFF_B[l] + o] for l, (_, o) in enumerate(SHAPES)]
    return W, b


def pack(W, b):
    p = np.zeros(N_PARAMS, dtype=np.float64)
    for l, (i, o) in enumerate(SHAPES):
        p[OFF_W[l]:OFF_W[l] + i * o] = np.asarray(W[l]).reshape(-1)
        p[OFF_B[l]:OFF_B[l] + o] = b[l]
    return p


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C0 * (z + GELU_C1 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C0 * (z + GELU_C1 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C0 * (1.0 + 3.0 * GELU_C1 * z * z)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def patches(images, idx, w, h):
    """Inputs of samples idx (global pixel indices over a stack of images
    [count][h][w][3]): the 3x3 clamp-edge neighbourhood, row-major over
    (dy, dx), channel fastest, minus 0.5; columns 27..31 zero."""
    images = np.asarray(images, dtype=np.float64).reshape(-1, h, w, 3)
    idx = np.asarray(idx, dtype=np.int64)
    n = w * h
    img, p = idx // n, idx % n
    y, x = p // w, p % w
    X = np.zeros((len(idx), IN_PAD))
    for nb in range(9):
        yy = np.clip(y + nb // 3 - 1, 0, h - 1)
        xx = np.clip(x + nb % 3 - 1, 0, w - 1)
        X[:, 3 * nb:3 * nb + 3] = images[img, yy, xx, :] - 0.5
    return X


def targets_of(targets, idx, layers, w, h, outputs):
    """Rows of the packed first passes [count][L][n][4] for samples idx."""
    n = w * h
    t = np.asarray(targets, dtype=np.float64).reshape(-1, layers, n, 4)
    idx = np.asarray(idx, dtype=np.int64)
    img, p = idx // n, idx % n
    T = np.zeros((len(idx), OUT_PAD))
    T[:, :outputs] = t[img, :, p, :].reshape(len(idx), layers * 4)[:, :outputs]
    return T


def forward(params, X, bf16_ops=False):
    """Returns (zs, hs, y): pre-activations, activations (hs[0] = X) and the
    sigmoid outputs.  bf16_ops rounds the product operands as the device
    does (inputs, weights and hidden activations)."""
    W, b = unpack(params)
    r = bf16 if bf16_ops else (lambda a: a)
    h = r(X)
    zs, hs = [], [h]
    for l in range(4):
        z = h @ r(W[l]) + b[l]
        zs.append(z)
        if l < 3:
            h = r(gelu(z))
            hs.append(h)
    return zs, hs, sigmoid(zs[3])


def loss(params, X, T, outputs):
    _, _, y = forward(params, X)
    e = (y - T)[:, :outputs]
    return float(np.sum(e * e) / (X.shape[0] * outputs))


# Derivative slips a backward pass could plausibly make, for tests that
# check that the device gradient's tolerance would catch them (loss_grad's
# `plant`): the GELU derivative without its second term, a hidden layer's
# error propagated through W instead of W^T, the sigmoid's derivative left
# out, the GELU derivative taken at the layer above's pre-activation.
PLANTS = ("gelu_first_term", "untransposed_weight", "no_sigmoid_grad", "gelu_wrong_layer")


def loss_grad(params, X, T, outputs, bf16_ops=False, plant=None):
    """Mean squared error over the valid outputs and its gradient with
    respect to the padded parameter vector.  plant (one of PLANTS) computes
    the gradient with that mistake, for tests of a tolerance's power."""
    W, _ = unpack(params)
    r = bf16 if bf16_ops else (lambda a: a)
    zs, hs, y = forward(params, X, bf16_ops)
    B = X.shape[0]
    mask = np.zeros(OUT_PAD)
    mask[:outputs] = 1.0
    e = (y - T) * mask
    L = float(np.sum(e * e) / (B * outputs))
    dsig = 1.0 if plant == "no_sigmoid_grad" else y * (1.0 - y)
    dz = 2.0 * e * dsig / (B * outputs)
    gW, gb = [None] * 4, [None] * 4
    for l in range(3, -1, -1):
        dzr = r(dz)
        gW[l] = hs[l].T @ dzr
        gb[l] = dz.sum(axis=0)
        if l > 0:
            Wl = r(W[l])
            back = dzr @ (Wl if (plant == "untransposed_weight" and Wl.shape[0] == Wl.shape[1]) else Wl.T)
            if plant == "gelu_first_term":
                t = np.tanh(GELU_C0 * (zs[l - 1] + GELU_C1 * zs[l - 1] ** 3))
                dz = back * (0.5 * (1.0 + t))
            elif plant == "gelu_wrong_layer" and zs[l].shape == zs[l - 1].shape:
                dz = back * gelu_grad(zs[l])
            else:
                dz = back * gelu_grad(zs[l - 1])
    return L, pack(gW, gb)


def adam(w, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 0); returns new (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1 - b1 ** (step + 1)
    c2 = 1 - b2 ** (step + 1)
    return w - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


_MASK = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _MASK
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _MASK
    return x ^ (x >> 31)


def sample_indices(seed, step, batch, pool):
    """Training batch of one step: sample b draws
    splitmix64(seed ^ splitmix64(step << 32 | b)) mod pool."""
    return np.array([splitmix64(seed ^ splitmix64(((step << 32) | b) & _MASK)) % pool for b in range(batch)],
                    dtype=np.int64)


def predict(params, rgb, layers, bf16_ops=True):
    """Packed first pass [layers][h][w][4] predicted for one image."""
    rgb = np.asarray(rgb, dtype=np.float64)
    h, w, _ = rgb.shape
    X = patches(rgb[None], np.arange(w * h), w, h)
    _, _, y = forward(params, X, bf16_ops)
    outputs = 4 * layers
    return y[:, :outputs].reshape(h * w, layers, 4).transpose(1, 0, 2).reshape(layers, h, w, 4)
