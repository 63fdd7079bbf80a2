"""The ladder surrogate's BF16 MFMA kernels against the CPU oracle
(oracle/mlp_oracle.py, which applies the same BF16 rounding points in
float64): GEMM lane layout, forward outputs, every parameter gradient, the
Adam update, and that device training lowers the loss."""
import os
import sys

import numpy as np
import pytest

from conftest import REPO

sys.path.insert(0, os.path.join(REPO, "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

DIMS = [516, 256, 256, 256, 64]  # the surrogate's shape


def say(*a):
    print(*a, flush=True)


def bits(x):
    """float64 values exactly representable in BF16 -> raw uint16 bits."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)
    return (u >> 16).astype(np.uint16)


@pytest.mark.parametrize("M_, N_, K_", [(32, 32, 8), (64, 96, 40), (96, 64, 544)])
def test_gemm_layout_exact(ptx, M_, N_, K_):
    say("gemm", M_, N_, K_)
    rng = np.random.default_rng(M_ + N_ + K_)
    a = rng.integers(-8, 9, size=(M_, K_)).astype(np.float64)
    b = rng.integers(-8, 9, size=(K_, N_)).astype(np.float64)
    c = ptx.gemm_bf16(bits(a), bits(b))
    assert np.array_equal(c.astype(np.float64), a @ b)  # integer sums are exact in FP32


def _setup(ptx, batch, seed):
    rng = np.random.default_rng(seed)
    net = ptx.Mlp(DIMS, max_batch=128, seed=seed)
    p = net.params().astype(np.float64)
    # non-zero biases so every epilogue term is exercised
    layers = M.unflatten(p, DIMS)
    layers = [(W, rng.normal(scale=0.05, size=b.shape)) for W, b in layers]
    net.set_params(M.flatten(layers).astype(np.float32))
    layers = M.unflatten(net.params().astype(np.float64), DIMS)
    x = rng.normal(size=(batch, DIMS[0])).astype(np.float32).astype(np.float64)
    return net, layers, x, rng


def test_forward_matches_oracle(ptx):
    net, layers, x, _ = _setup(ptx, 50, 5)
    y = net.forward(x.astype(np.float32)).astype(np.float64)
    yo = M.forward(layers, x, emulate=True)[2]
    err = np.max(np.abs(y - yo) / (np.abs(yo) + 1e-3))
    say("forward max rel err", err)
    assert err < 2e-3


def test_gradients_match_oracle(ptx):
    net, layers, x, rng = _setup(ptx, 50, 6)
    t = rng.uniform(0.3, 3.0, size=(50, DIMS[-1]))
    mode = rng.integers(0, 3, size=(50, DIMS[-1])).astype(np.uint8)
    loss = net.loss_grad(x.astype(np.float32), t.astype(np.float32), mode)
    lo, _, go = M.loss_grad(layers, x, t.astype(np.float32).astype(np.float64), mode, emulate=True)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    g = M.unflatten(net.grads().astype(np.float64), DIMS)
    for l, ((dW, db), (oW, ob)) in enumerate(zip(g, go)):
        eW = np.linalg.norm(dW - oW) / np.linalg.norm(oW)
        eb = np.linalg.norm(db - ob) / np.linalg.norm(ob)
        say(f"layer {l}: rel err dW {eW:.2e} db {eb:.2e}")
        assert eW < 2e-3 and eb < 2e-3


def test_adam_update(ptx):
    net, layers, x, rng = _setup(ptx, 40, 7)
    t = rng.uniform(0.3, 3.0, size=(40, DIMS[-1])).astype(np.float32)
    mode = np.ones((40, DIMS[-1]), np.uint8)
    p0 = net.params().astype(np.float64)
    net.loss_grad(x.astype(np.float32), t, mode)
    g = net.grads().astype(np.float64)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    net.adam(lr, b1, b2, eps)
    p1 = net.params().astype(np.float64)
    m = (1 - b1) * g
    v = (1 - b2) * g * g
    ref = p0 - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    assert np.allclose(p1, ref, rtol=1e-5, atol=1e-7)


def test_training_lowers_loss(ptx):
    rng = np.random.default_rng(8)
    n = 512
    x = rng.normal(size=(n, DIMS[0])).astype(np.float32)
    proj = rng.normal(size=(DIMS[0], DIMS[-1])) / np.sqrt(DIMS[0])
    t = (np.log1p(np.exp(x @ proj)) + 0.2).astype(np.float32)
    mode = np.ones_like(t, dtype=np.uint8)
    net = ptx.Mlp(DIMS, max_batch=128, seed=3)
    l0 = net.loss_grad(x[:128], t[:128], mode[:128])
    losses = net.train(x, t, mode, batch=128, epochs=40, lr=1e-3, seed=1)
    say("loss", l0, "->", losses[-1])
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.2 * l0
