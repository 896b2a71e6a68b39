"""The warm-start network's kernels against the CPU oracle (oracle/mlp_oracle.py).

* feature map: float32 device values against the float64 map;
* loss and gradient of one forward/backward pass (no update): against the
  oracle's bfloat16-rounding form, which rounds where the device stores
  bfloat16 and sums in float64 (the device sums in float32, so single
  roundings may flip); the tolerance is per layer, relative L2 of the weight
  and bias gradients, and the oracle's form is itself pinned to finite
  differences of the exact loss on the CPU;
* Adam: the device's update against the oracle's, applied to the device's
  own gradient, over two steps (moments carried);
* prediction against the oracle's forward, in one batch and in chunks;
* training on a smooth synthetic target lowers the loss, and two runs with
  the same seed are bit-identical.
"""
import numpy as np
import pytest

from fvs2d_amd.mlp import WarmStartNet
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _raw(n, seed=0):
    rng = np.random.default_rng(seed)
    r = np.exp(rng.uniform(np.log(0.01), np.log(50), n))
    th = rng.uniform(-np.pi, np.pi, n)
    return np.stack([0.5 + r * np.cos(th), r * np.sin(th), rng.uniform(0, 50, n), rng.uniform(0.5, 0.85, n),
                     rng.uniform(-2, 4, n), rng.uniform(0.1, 0.15, n)], 1).astype(np.float32)


def _target(raw):
    x, y, d, M, a, t = raw.astype(np.float64).T
    r = np.hypot(x - 0.5, y)
    th = np.arctan2(y, x - 0.5)
    return np.stack([np.sin(2 * th) * np.exp(-r), np.cos(th) / (1 + r), (M - 0.675) * 2, a / 4 + t], 1)


def _layer_slices(depth):
    out, q = [], 0
    for i, o in mo.layer_sizes(depth):
        out.append((slice(q, q + i * o), slice(q + i * o, q + i * o + o)))
        q += i * o + o
    return out


def test_features_match_oracle():
    net = WarmStartNet(depth=1, batch=1024, seed=3)
    raw = _raw(1000)
    f = net.features(raw)
    ref = mo.features(raw.astype(np.float64))
    assert np.max(np.abs(f - ref)) < 2e-6


@pytest.mark.parametrize("depth,n", [(2, 200), (5, 512)])
def test_gradient_matches_oracle(depth, n):
    net = WarmStartNet(depth=depth, batch=512, seed=depth)
    p = net.get_params().astype(np.float64)
    assert p.size == mo.nparams(depth)
    raw = _raw(n, seed=depth)
    T = _target(raw).astype(np.float32)
    loss, g = net.grad(raw, T)
    X = mo.features(raw.astype(np.float64)).astype(np.float32).astype(np.float64)
    l_ref, g_ref = mo.loss_and_grad(p, X, T.astype(np.float64), depth, bf16_mode=True)
    assert abs(loss - l_ref) <= 1e-3 * l_ref
    for k, (sw, sb) in enumerate(_layer_slices(depth)):
        for s in (sw, sb):
            a, b = g[s].astype(np.float64), g_ref[s]
            assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-12, (k, s)
    # and the exact (float64, no rounding) gradient of the same parameters
    _, g_exact = mo.loss_and_grad(p, X, T.astype(np.float64), depth)
    assert np.linalg.norm(g - g_exact) <= 6e-2 * np.linalg.norm(g_exact)


def test_adam_matches_oracle():
    depth = 2
    net = WarmStartNet(depth=depth, batch=256, lr=1e-3, seed=11)
    raw = _raw(256, seed=5)
    T = _target(raw).astype(np.float32)
    p = net.get_params().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in (1, 2):
        _, g = net.grad(raw, T)
        net.step(raw, T)
        p, m, v = mo.adam(p, g.astype(np.float64), m, v, t, lr=1e-3)
        got = net.get_params().astype(np.float64)
        assert np.max(np.abs(got - p) / (np.abs(p) + 1e-3)) < 2e-5, t
        p = got


def test_predict_matches_forward():
    depth = 3
    net = WarmStartNet(depth=depth, batch=256, seed=7)
    raw = _raw(700, seed=9)
    y = net.predict(raw)                      # three chunks of at most 256 rows
    p = net.get_params().astype(np.float64)
    X = mo.features(raw.astype(np.float64)).astype(np.float32).astype(np.float64)
    y_ref, _ = mo.forward(p, X, depth, bf16_mode=True)
    assert np.max(np.abs(y - y_ref)) <= 1e-2 * np.max(np.abs(y_ref))
    assert np.array_equal(net.predict(raw[256:512]), y[256:512])


def test_training_lowers_loss_and_is_deterministic():
    raw = _raw(20000, seed=1)
    T = _target(raw).astype(np.float32)
    params = []
    for _ in range(2):
        net = WarmStartNet(depth=3, batch=4096, lr=1e-3, seed=21)
        net.set_data(raw, T)
        loss = net.train(300)
        assert np.all(np.isfinite(loss))
        assert loss[-20:].mean() < 0.1 * loss[:5].mean()
        params.append(net.get_params())
    assert np.array_equal(params[0], params[1])
