"""CPU oracle of the warm-start network (test infrastructure only).

The reference has no learned part: this network, its feature map and its
training are this library's own (the warm start of steady solves).  The
oracle restates, in numpy float64, what the device code computes:

* ``features``: the 68-entry map of a cell record (x, y, wall distance,
  Mach, angle of attack in degrees, thickness);
* ``forward`` / ``loss_and_grad``: the network 68 -> depth x 256 (GELU) -> 4,
  mean-square loss, and its gradient by hand-written back-propagation;
* ``adam``: one Adam update (Kingma & Ba 2015, bias-corrected).

With ``bf16=True`` the forward and backward round to bfloat16 (round to
nearest even) exactly where the device stores bfloat16: the network input,
the weights of the hidden layers, the pre-activations Z and activations H,
and the back-propagated dZ; all sums are float64 here and float32 on the
device.  Without it the arithmetic is plain float64, which is what the
finite-difference check pins.

Only tests and the benchmark's CPU leg use this module.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

WIDTH = 256
OUT = 4
NFEAT = 68
NRAW = 6
XC = 0.5            # mid-chord: origin of (r, theta)
R_SCALE = 4.0       # s = ln(1 + 4 r) / ln(1 + 4 * 50)
R_FAR = 50.0
D_SCALE = 100.0     # wall-distance feature ln(1 + 100 d) / ln(1 + 100 * 50)
NFREQ = 16
COND_SHIFT = (0.675, 1.0, 0.125)   # Mach, AoA (deg), thickness
COND_SCALE = (0.175, 3.0, 0.025)


def features(raw) -> np.ndarray:
    """Feature map of cell records raw[n, 6] = (x, y, d, M, aoa_deg, t):
    [sin(k pi s), cos(k pi s)] for k = 1..16, then [sin(k th), cos(k th)] for
    k = 1..16, then the scaled wall distance, Mach, AoA and thickness."""
    raw = np.asarray(raw, np.float64)
    x, y, d, M, a, t = raw.T
    dx = x - XC
    r = np.hypot(dx, y)
    th = np.arctan2(y, dx)
    s = np.log1p(R_SCALE * r) / math.log1p(R_SCALE * R_FAR)
    k = np.arange(1, NFREQ + 1)
    f = np.empty((raw.shape[0], NFEAT))
    f[:, 0:32:2] = np.sin(np.pi * s[:, None] * k)
    f[:, 1:32:2] = np.cos(np.pi * s[:, None] * k)
    f[:, 32:64:2] = np.sin(th[:, None] * k)
    f[:, 33:64:2] = np.cos(th[:, None] * k)
    f[:, 64] = np.log1p(D_SCALE * d) / math.log1p(D_SCALE * R_FAR)
    f[:, 65] = (M - COND_SHIFT[0]) / COND_SCALE[0]
    f[:, 66] = (a - COND_SHIFT[1]) / COND_SCALE[1]
    f[:, 67] = (t - COND_SHIFT[2]) / COND_SCALE[2]
    return f


def layer_sizes(depth: int):
    return [(NFEAT, WIDTH)] + [(WIDTH, WIDTH)] * (depth - 1) + [(WIDTH, OUT)]


def nparams(depth: int) -> int:
    return sum(i * o + o for i, o in layer_sizes(depth))


def unflatten(p, depth: int):
    """Flat parameter vector -> [(W [in, out], b [out]), ...]: per layer the
    weights row by row (input index slow), then the biases."""
    out, q = [], 0
    p = np.asarray(p, np.float64)
    for i, o in layer_sizes(depth):
        W = p[q:q + i * o].reshape(i, o)
        q += i * o
        b = p[q:q + o]
        q += o
        out.append((W, b))
    assert q == p.size
    return out


def flatten(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def bf16(x) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, X, depth: int, bf16_mode: bool = False):
    """Network output y[n, 4] and the per-layer (input, Z) cache."""
    rnd = bf16 if bf16_mode else (lambda v: v)
    layers = unflatten(p, depth)
    h = rnd(X)
    cache = []
    for W, b in layers[:-1]:
        z = h @ rnd(W) + b
        cache.append((h, rnd(z)))
        h = rnd(gelu(z))
    W, b = layers[-1]
    y = h @ W + b
    cache.append((h, None))
    return y, cache


def loss_and_grad(p, X, T, depth: int, bf16_mode: bool = False):
    """Mean-square loss sum((y - T)^2) / (4 n) and its gradient (flat, the
    layout of ``unflatten``)."""
    rnd = bf16 if bf16_mode else (lambda v: v)
    n = X.shape[0]
    y, cache = forward(p, X, depth, bf16_mode)
    r = y - T
    loss = float(np.sum(r * r) / (OUT * n))
    dy = 2.0 * r / (OUT * n)
    layers = unflatten(p, depth)
    grads = [None] * len(layers)
    h, _ = cache[-1]
    W, _ = layers[-1]
    grads[-1] = (h.T @ dy, dy.sum(0))
    dz = rnd((dy @ W.T) * gelu_grad(cache[-2][1]))
    for l in range(len(layers) - 2, -1, -1):
        h, _ = cache[l]
        grads[l] = (h.T @ dz, dz.sum(0))
        if l > 0:
            W, _ = layers[l]
            dz = rnd((dz @ rnd(W).T) * gelu_grad(cache[l - 1][1]))
    return loss, flatten(grads)


def adam(p, g, m, v, t: int, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (t counts from 1).  Returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
