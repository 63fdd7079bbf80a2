"""The CPU restatement of the ladder-surrogate network (oracle/mlp_oracle.py)
pinned without the GPU: BF16 rounding against exactly representable cases,
and its analytic gradient against central finite differences of the float64
network (so a dropped term, a wrong transpose or a wrong mask fails)."""
import os
import sys

import numpy as np

from conftest import REPO

sys.path.insert(0, os.path.join(REPO, "oracle"))
import mlp_oracle as M  # noqa: E402


def test_bf16_rounding_known_values():
    # BF16 keeps 8 significant bits: 1 + k/128.
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1.0 + 1 / 256) == 1.0            # tie -> even (1.0)
    assert M.bf16(1.0 + 3 / 256) == 1.0 + 2 / 128  # tie -> even (k = 2)
    assert M.bf16(1.0 + 1 / 256 + 1 / 4096) == 1.0 + 1 / 128  # above the tie rounds up
    assert M.bf16(-3.0) == -3.0
    assert M.bf16(2.0 ** -20) == 2.0 ** -20
    x = np.random.default_rng(0).normal(size=1000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-300)
    assert np.array_equal(M.bf16(r), r)  # idempotent


def _net(dims, rng):
    return [(rng.normal(scale=1 / np.sqrt(a), size=(a, b)), rng.normal(scale=0.1, size=b))
            for a, b in zip(dims[:-1], dims[1:])]


def test_flatten_roundtrip():
    rng = np.random.default_rng(1)
    dims = [5, 7, 3]
    layers = _net(dims, rng)
    flat = M.flatten(layers)
    assert len(flat) == 5 * 7 + 7 + 7 * 3 + 3
    back = M.unflatten(flat, dims)
    for (W, b), (W2, b2) in zip(layers, back):
        assert np.array_equal(W, W2) and np.array_equal(b, b2)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    dims = [6, 9, 8, 5]
    layers = _net(dims, rng)
    B = 7
    x = rng.normal(size=(B, dims[0]))
    t = rng.uniform(0.2, 2.0, size=(B, dims[-1]))
    mode = rng.integers(0, 3, size=(B, dims[-1])).astype(np.uint8)
    _, y, grads = M.loss_grad(layers, x, t, mode, emulate=False)
    # keep the one-sided penalty away from its kink and ReLUs away from 0
    t = np.where((mode == 2) & (np.abs(t - y) < 0.05), y + 0.3, t)
    _, _, grads = M.loss_grad(layers, x, t, mode, emulate=False)
    flat = M.flatten(layers)
    g = M.flatten(grads)
    h = 1e-6
    idx = rng.choice(len(flat), size=60, replace=False)
    for i in idx:
        fp = flat.copy()
        fm = flat.copy()
        fp[i] += h
        fm[i] -= h
        lp = M.loss_grad(M.unflatten(fp, dims), x, t, mode, emulate=False)[0]
        lm = M.loss_grad(M.unflatten(fm, dims), x, t, mode, emulate=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_emulated_network_close_to_exact():
    rng = np.random.default_rng(3)
    dims = [16, 32, 8]
    layers = _net(dims, rng)
    x = rng.normal(size=(4, 16))
    y0 = M.forward(layers, x, emulate=False)[2]
    y1 = M.forward(layers, x, emulate=True)[2]
    assert np.max(np.abs(y0 - y1)) < 0.05 * np.max(np.abs(y0))
    assert not np.array_equal(y0, y1)  # the roundings are applied
