"""Deterministic instance generators for the benchmark workloads.

Every generator is a pure function of its seed (numpy's Philox bit generator)
and returns an undirected COO edge list (src, dst, w) with 0-based indices,
which can be written as a gset file (0-based, so the file name should contain
".gset", the reader rule of src/problem.cc:208).

Sign convention: the reference's energy is E = sum_{i<j} w_ij s_i s_j
(src/problem.cc:121-142), so a +-J spin glass with couplings J_ij in the usual
-sum J s s convention is written with file weight w = -J.
"""
import numpy as np


def _rng(seed, stream=0):
    return np.random.Generator(np.random.Philox(key=[int(seed), int(stream)]))


def torus_2d(L, seed=0, pm=True):
    """Periodic L x L lattice; pm: w = +-1 with probability 1/2, else w = -1."""
    idx = np.arange(L * L).reshape(L, L)
    right = np.stack([idx.ravel(), np.roll(idx, -1, axis=1).ravel()], 1)
    down = np.stack([idx.ravel(), np.roll(idx, -1, axis=0).ravel()], 1)
    e = np.concatenate([right, down])
    if pm:
        w = np.where(_rng(seed, 2).integers(0, 2, len(e)) == 1, 1.0, -1.0)
    else:
        w = -np.ones(len(e))
    return L * L, e[:, 0].astype(np.int32), e[:, 1].astype(np.int32), w


def ea_3d(L, seed=0):
    """Periodic L^3 Edwards-Anderson +-J lattice (3 bonds per site)."""
    idx = np.arange(L ** 3).reshape(L, L, L)
    parts = [np.stack([idx.ravel(), np.roll(idx, -1, axis=a).ravel()], 1) for a in range(3)]
    e = np.concatenate(parts)
    w = np.where(_rng(seed, 3).integers(0, 2, len(e)) == 1, 1.0, -1.0)
    return L ** 3, e[:, 0].astype(np.int32), e[:, 1].astype(np.int32), w


def erdos_renyi(n, p, seed=0, pm=False):
    """G(n, p) without self-loops; pm: w = +-1, else w = +1 (MaxCut, J = -w)."""
    g = _rng(seed, 4)
    m = g.binomial(n * (n - 1) // 2, p)
    # sample m distinct unordered pairs: batches of candidate pairs, taken in
    # draw order, skipping self-loops and pairs already chosen
    chosen = np.zeros(0, np.int64)
    while len(chosen) < m:
        need = m - len(chosen)
        u = g.integers(0, n, 2 * need + 16)
        v = g.integers(0, n, 2 * need + 16)
        key = np.minimum(u, v).astype(np.int64) * n + np.maximum(u, v)
        ok = u != v
        _, first = np.unique(key, return_index=True)
        is_first = np.zeros(len(key), bool)
        is_first[first] = True
        ok &= is_first & ~np.isin(key, chosen)
        chosen = np.concatenate([chosen, key[ok][:need]])
    chosen = np.sort(chosen)
    out = np.stack([chosen // n, chosen % n], 1)
    e = out.astype(np.int32).reshape(-1, 2)
    if pm:
        w = np.where(g.integers(0, 2, len(e)) == 1, 1.0, -1.0)
    else:
        w = np.ones(len(e))
    return n, e[:, 0], e[:, 1], w


def ea_batch(n_inst, seed=0, sizes=(6, 8, 10, 12)):
    """The batched workload: n_inst 3D EA +-J instances, L drawn uniformly
    from `sizes`; instance i uses seed 1000 + i.  Returns (Ls, [coo, ...])."""
    Ls = np.random.default_rng(seed).choice(list(sizes), size=n_inst)
    return Ls, [ea_3d(int(L), seed=1000 + i) for i, L in enumerate(Ls)]


def write_gset(path, n, src, dst, w):
    with open(path, "w") as f:
        f.write(f"{n} {len(src)}\n")
        for a, b, c in zip(src.tolist(), dst.tolist(), w.tolist()):
            f.write(f"{a} {b} {int(c)}\n")
