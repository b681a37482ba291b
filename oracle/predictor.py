"""CPU oracle of the learned initial-iterate predictor -- TEST INFRASTRUCTURE.

Only tests/ (and bench.py's checks) import this module, and only as the
checker of the HIP kernels in channel_flow_amd/csrc/cf_predictor.hip.  The
network has no counterpart in the reference (whose physical step starts at
U^n, main.cpp:444-464); it is defined here, in FP64 NumPy, forward and
backward written out by hand.

Network (per cell; 16 input features x', 64 hidden channels, 5 outputs):

    h1(q)  = W1 x'(q)                      pointwise 16 -> 64 (x'[15] is the
                                           in-domain mask, so W1[:, 15] is the
                                           bias and h1 = 0 outside the domain)
    z2(p)  = b2 + sum_t W2[t] h1(p + o_t)  27-point 3x3x3 conv 64 -> 64
    a(p)   = gelu(z2(p))                   0.5 z (1 + erf(z / sqrt 2))
    y(p)   = b3 + W3 a(p)                  pointwise 64 -> 5

with tap t = 9 (di+1) + 3 (dj+1) + (dk+1), o_t = (di, dj, dk).  Volumes are
[n][I][J][K][C] with the conv taking zeros outside the volume; outputs are
formed on an inner region (margin 1 on every side for training crops).

Loss: L = sum over inner cells and the 5 outputs of (y - target)^2 / (5 N).

bf16=True rounds, as the HIP kernels do, x', W1, W2 and h1 to bfloat16
(round to nearest even) before they are used; sums stay FP64 here (FP32 on
the GPU).
"""
from __future__ import annotations

import math

import numpy as np

C_IN = 16
C_HID = 64
C_OUT = 5
SHAPES = {"W1": (C_HID, C_IN), "W2": (27, C_HID, C_HID), "b2": (C_HID,), "W3": (C_OUT, C_HID), "b3": (C_OUT,)}
ORDER = ["W1", "W2", "b2", "W3", "b3"]


def n_params() -> int:
    return sum(int(np.prod(SHAPES[k])) for k in ORDER)


def unpack(flat: np.ndarray) -> dict:
    """Flat parameter vector (the order of ORDER, each C-order) -> dict."""
    out, o = {}, 0
    for k in ORDER:
        n = int(np.prod(SHAPES[k]))
        out[k] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(SHAPES[k])
        o += n
    return out


def pack(p: dict) -> np.ndarray:
    return np.concatenate([np.asarray(p[k], dtype=np.float64).ravel() for k in ORDER])


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _erf(x):
    from scipy.special import erf
    return erf(x)


def taps():
    return [(di, dj, dk) for di in (-1, 0, 1) for dj in (-1, 0, 1) for dk in (-1, 0, 1)]


def _shift(v, o, inner, margin):
    """v[:, margin + o + i] over the inner region (zeros outside v)."""
    n, I, J, K, C = v.shape
    pad = np.zeros((n, I + 2, J + 2, K + 2, C))
    pad[:, 1:-1, 1:-1, 1:-1] = v
    (a0, a1), (b0, b1), (c0, c1) = inner
    di, dj, dk = o
    return pad[:, 1 + a0 + di:1 + a1 + di, 1 + b0 + dj:1 + b1 + dj, 1 + c0 + dk:1 + c1 + dk]


def forward(p: dict, x: np.ndarray, margin: int = 1, bf: bool = False):
    """x: [n][I][J][K][16].  Returns y on the inner region (margin cells
    dropped on every side) and the intermediates the backward pass needs."""
    n, I, J, K, _ = x.shape
    inner = ((margin, I - margin), (margin, J - margin), (margin, K - margin))
    W1, W2 = p["W1"], p["W2"]
    if bf:
        x, W1, W2 = bf16(x), bf16(W1), bf16(W2)
    h1 = x @ W1.T
    if bf:
        h1 = bf16(h1)
    z2 = np.zeros(tuple(b - a for a, b in inner) + (C_HID,))
    z2 = np.broadcast_to(p["b2"], (n,) + z2.shape[:3] + (C_HID,)).copy()
    for t, o in enumerate(taps()):
        z2 += _shift(h1, o, inner, margin) @ W2[t].T
    a = gelu(z2)
    y = a @ p["W3"].T + p["b3"]
    return y, dict(x=x, h1=h1, z2=z2, a=a, inner=inner, W1=W1, W2=W2)


def loss(y, target):
    return float(np.sum((y - target) ** 2) / (C_OUT * y[..., 0].size))


def backward(p: dict, cache: dict, y: np.ndarray, target: np.ndarray) -> dict:
    """Gradients of loss(y, target) with respect to every parameter."""
    x, h1, z2, a, inner = cache["x"], cache["h1"], cache["z2"], cache["a"], cache["inner"]
    N = y[..., 0].size
    dy = 2.0 * (y - target) / (C_OUT * N)
    g = {}
    g["W3"] = dy.reshape(-1, C_OUT).T @ a.reshape(-1, C_HID)
    g["b3"] = dy.reshape(-1, C_OUT).sum(axis=0)
    dz2 = (dy @ p["W3"]) * gelu_grad(z2)
    g["b2"] = dz2.reshape(-1, C_HID).sum(axis=0)
    n, I, J, K, _ = x.shape
    W2 = cache["W2"]
    gW2 = np.zeros(SHAPES["W2"])
    dh1 = np.zeros(h1.shape)
    (a0, a1), (b0, b1), (c0, c1) = inner
    for t, (di, dj, dk) in enumerate(taps()):
        hs = _shift(h1, (di, dj, dk), inner, 1)
        gW2[t] = dz2.reshape(-1, C_HID).T @ hs.reshape(-1, C_HID)
        # scatter dz2 W2[t] back to the cells the tap read (inside the volume)
        contrib = dz2 @ W2[t]
        pad = np.zeros((n, I + 2, J + 2, K + 2, C_HID))
        pad[:, 1 + a0 + di:1 + a1 + di, 1 + b0 + dj:1 + b1 + dj, 1 + c0 + dk:1 + c1 + dk] += contrib
        dh1 += pad[:, 1:-1, 1:-1, 1:-1]
    g["W2"] = gW2
    g["W1"] = dh1.reshape(-1, C_HID).T @ x.reshape(-1, C_IN)
    return g


def init_params(seed: int) -> dict:
    """He-style initial weights; the last layer starts at zero, so the
    untrained network predicts zero (the base guess)."""
    rng = np.random.default_rng(seed)
    p = {"W1": rng.standard_normal(SHAPES["W1"]) * math.sqrt(1.0 / C_IN),
         "W2": rng.standard_normal(SHAPES["W2"]) * math.sqrt(2.0 / (27 * C_HID)),
         "b2": np.zeros(SHAPES["b2"]),
         "W3": np.zeros(SHAPES["W3"]),
         "b3": np.zeros(SHAPES["b3"])}
    return p
