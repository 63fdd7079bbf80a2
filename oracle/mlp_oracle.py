"""CPU restatement of the ladder-surrogate network (test infrastructure only).

Loaded only by tests/ as the checker of the GPU training and inference
kernels; the product never imports it.  Float64 numpy throughout.  With
``emulate=True`` it applies the BF16 roundings at exactly the points the
GPU kernels do (operands of every GEMM rounded to BF16, round to nearest
even; biases, pre-activations and gradients kept wide), so the remaining
difference to the device is FP32 accumulation order.  With ``emulate=False``
it is the exact float64 network, whose gradient the tests pin by central
finite differences.

Network: x -> [Z_l = X_l W_l + b_l, X_{l+1} = relu(Z_l)] -> o = Z_last,
y = softplus(o).  Loss per output element: mode 1 (y - t)^2, mode 2
max(0, t - y)^2, mode 0 nothing; total divided by the batch size.
Flat parameter layout: per layer W[in][out] row-major, then b[out].

There is no counterpart in the reference (it computes the ladder with
get_beta, src/cae_estimator.cc:295-428); the network's targets are that
function's ladders.
"""
import numpy as np


def bf16(x):
    """Round to BF16 (nearest even) through float32, returned as float64."""
    f = np.asarray(x, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unflatten(flat, dims):
    out, o = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = np.asarray(flat[o:o + a * b], dtype=np.float64).reshape(a, b)
        o += a * b
        bb = np.asarray(flat[o:o + b], dtype=np.float64)
        o += b
        out.append((W, bb))
    assert o == len(flat)
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def softplus(o):
    return np.maximum(o, 0.0) + np.log1p(np.exp(-np.abs(o)))


def sigmoid(o):
    e = np.exp(-np.abs(o))
    return np.where(o >= 0, 1.0 / (1.0 + e), e / (1.0 + e))


def forward(layers, x, emulate=True):
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    X = [r(x)]
    Z = []
    for l, (W, b) in enumerate(layers):
        z = X[-1] @ r(W) + b
        Z.append(z)
        if l + 1 < len(layers):
            X.append(r(np.maximum(z, 0.0)))
    return X, Z, softplus(Z[-1])


def loss_grad(layers, x, t, mode, emulate=True):
    """Returns (loss, y, [(dW, db), ...])."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    B = x.shape[0]
    X, Z, y = forward(layers, x, emulate)
    o = Z[-1]
    d1 = np.where(mode == 1, y - t, 0.0)
    d2 = np.where(mode == 2, np.maximum(t - y, 0.0), 0.0)
    loss = (np.sum(d1 * d1) + np.sum(d2 * d2)) / B
    dy = 2.0 * d1 - 2.0 * d2
    G = dy * sigmoid(o) / B
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        Gb = r(G)
        grads[l] = (X[l].T @ Gb, G.sum(axis=0))
        if l > 0:
            G = (Gb @ r(W).T) * (Z[l - 1] > 0)
    return loss, y, grads
