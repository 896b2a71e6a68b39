"""Pins of the warm-start network's CPU oracle (oracle/mlp_oracle.py).

* GELU against math.erf point by point, its derivative against central
  differences;
* the feature map against identities it must satisfy (unit sin/cos pairs,
  the double-angle relation between harmonics 1 and 2, recovery of theta
  and s, the scalar channels at known points);
* the hand-written gradient against central finite differences of the
  float64 loss, for parameters of every layer of a depth-2 network;
* bfloat16 rounding at known values (ties to even);
* Adam against its closed form for a constant gradient (every step moves a
  parameter by lr * g / (|g| + eps));
* the bfloat16-rounding gradient stays close to the float64 one, so the
  device comparison (which uses the rounding form) tests the same function.
"""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_gelu_pointwise_and_derivative():
    z = np.linspace(-6, 6, 121)
    ref = np.array([0.5 * v * (1 + math.erf(v / math.sqrt(2))) for v in z])
    assert np.max(np.abs(mo.gelu(z) - ref)) < 1e-15
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(mo.gelu_grad(z) - fd)) < 1e-8
    assert mo.gelu(np.array([0.0]))[0] == 0.0 and abs(mo.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15


def _raw(n, seed=0):
    rng = np.random.default_rng(seed)
    r = np.exp(rng.uniform(np.log(0.01), np.log(50), n))
    th = rng.uniform(-np.pi, np.pi, n)
    return np.stack([0.5 + r * np.cos(th), r * np.sin(th), rng.uniform(0, 50, n), rng.uniform(0.5, 0.85, n),
                     rng.uniform(-2, 4, n), rng.uniform(0.1, 0.15, n)], 1)


def test_feature_identities():
    raw = _raw(500)
    f = mo.features(raw)
    assert f.shape == (500, 68)
    for j in range(0, 64, 2):
        assert np.allclose(f[:, j] ** 2 + f[:, j + 1] ** 2, 1.0, atol=1e-14)
    # harmonic 2 from harmonic 1 (both families)
    for base in (0, 32):
        s1, c1, s2, c2 = f[:, base], f[:, base + 1], f[:, base + 2], f[:, base + 3]
        assert np.allclose(s2, 2 * s1 * c1, atol=1e-13) and np.allclose(c2, c1 * c1 - s1 * s1, atol=1e-13)
    th = np.arctan2(raw[:, 1], raw[:, 0] - 0.5)
    assert np.allclose(np.arctan2(f[:, 32], f[:, 33]), th, atol=1e-12)
    r = np.hypot(raw[:, 0] - 0.5, raw[:, 1])
    s = np.log1p(4 * r) / np.log(201.0)
    assert np.allclose(np.arccos(np.clip(f[:, 1], -1, 1)) / np.pi, np.minimum(s, 2 - s) % 2, atol=1e-6)
    # scalar channels at known points
    one = mo.features(np.array([[50.5, 0.0, 50.0, 0.675, 1.0, 0.125]]))[0]
    assert abs(one[64] - 1.0) < 1e-15 and np.allclose(one[65:], 0.0, atol=1e-15)
    assert abs(one[0]) < 1e-12 and abs(one[1] + 1.0) < 1e-12   # s = 1: sin(pi) = 0, cos(pi) = -1
    assert abs(one[32]) < 1e-15 and abs(one[33] - 1.0) < 1e-15  # theta = 0
    hi = mo.features(np.array([[0.5, 1.0, 0.0, 0.85, 4.0, 0.15]]))[0]
    assert hi[64] == 0.0 and np.allclose(hi[65:], [1.0, 1.0, 1.0])
    assert abs(hi[32] - 1.0) < 1e-15 and abs(hi[33]) < 1e-15    # theta = pi/2


def test_gradient_against_finite_differences():
    depth = 2
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.1, mo.nparams(depth))
    X = rng.normal(0, 1, (16, 68))
    T = rng.normal(0, 1, (16, 4))
    loss, g = mo.loss_and_grad(p, X, T, depth)
    # loss value from the forward alone
    y, _ = mo.forward(p, X, depth)
    assert abs(loss - np.mean((y - T) ** 2)) < 1e-14
    offs = np.cumsum([0] + [i * o + o for i, o in mo.layer_sizes(depth)])
    picks = np.concatenate([rng.integers(a, b, 12) for a, b in zip(offs[:-1], offs[1:])])
    h = 1e-6
    for q in picks:
        e = np.zeros_like(p)
        e[q] = h
        fd = (mo.loss_and_grad(p + e, X, T, depth)[0] - mo.loss_and_grad(p - e, X, T, depth)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(g[q])) + 1e-9, (q, fd, g[q])


def test_bf16_rounding():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e38, 1e-40])
    r = mo.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0                      # tie -> even
    assert r[2] == 1 + 2 ** -6                              # tie -> even (up)
    assert r[3] == 1 + 2 ** -7 and r[4] == -2.5
    assert abs(r[5] / 3.0e38 - 1) < 2 ** -8
    assert abs(r[6] - 1e-40) <= 2.0 ** -134                  # subnormal spacing 2^-133


def test_adam_closed_form():
    p0 = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -1e-3, 2.0])
    m = v = np.zeros(3)
    p = p0
    for t in range(1, 6):
        p, m, v = mo.adam(p, g, m, v, t, lr=1e-2)
        step = 1e-2 * g / (np.abs(g) + 1e-8)
        assert np.allclose(p, p0 - t * step, rtol=0, atol=1e-12)


@pytest.mark.parametrize("depth", [1, 3])
def test_bf16_form_close_to_exact(depth):
    rng = np.random.default_rng(2)
    sizes = mo.layer_sizes(depth)
    layers = [(rng.uniform(-1, 1, (i, o)) * math.sqrt(6 / (i + o)), rng.normal(0, 0.05, o)) for i, o in sizes]
    p = mo.flatten(layers)
    X = mo.features(_raw(256, 3))
    T = rng.normal(0, 0.3, (256, 4))
    l0, g0 = mo.loss_and_grad(p, X, T, depth)
    l1, g1 = mo.loss_and_grad(p, X, T, depth, bf16_mode=True)
    assert abs(l1 - l0) < 2e-2 * l0
    assert np.linalg.norm(g1 - g0) < 5e-2 * np.linalg.norm(g0)
    assert mo.nparams(5) == 281860
