"""Host-side logic of the ladder surrogate (no GPU): label encoding of a
multihistogram ladder and its inverse, and the pilot-histogram features."""
import os
import sys

import numpy as np

from conftest import REPO

sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
import surrogate as S  # noqa: E402


def test_labels_roundtrip_to_ladder():
    bs, be = 2.0, 1.0 / 3.0
    rng = np.random.default_rng(0)
    for k in (2, 5, 23, 40, 64):
        inc = rng.uniform(0.5, 1.5, size=k - 1)
        lad = bs - np.concatenate([[0.0], np.cumsum(inc)]) * (bs - be) / 70.0
        assert lad[-1] > be
        t, m = S.labels(lad, bs, be)
        assert m[: k - 1].tolist() == [1] * (k - 1)
        if k < S.N_OUT:
            assert m[k - 1] == 2 and np.all(m[k:] == 0)
            y = t.astype(np.float64).copy()
            y[k - 1] += 1e-3  # just past the hinge: the next step crosses beta_end
        else:
            y = np.concatenate([t, [0.0]])
        out = S.ladder_from_increments(y, bs, be)
        assert len(out) == k
        assert np.allclose(out, lad, rtol=0, atol=1e-6)


def test_ladder_from_increments_stops_before_beta_end():
    out = S.ladder_from_increments(np.full(64, 4.0), 2.0, 1.0)  # s = 1/32, 4 s = 0.125 per step
    assert np.all(out > 1.0) and len(out) == 8
    assert len(S.ladder_from_increments(np.full(64, 1e-3), 2.0, 1.0)) == S.N_OUT


def test_features_shape_and_normalisation():
    nb = 200
    counts = np.zeros((8, nb), np.uint32)
    rng = np.random.default_rng(1)
    for k in range(8):
        c = 40 + 15 * k
        counts[k, c - 10:c + 10] = rng.integers(1, 50, 20)
    stats = {"n": 400, "mean_degree": 4.0, "pm": 1.0}
    f = S.features(counts, -400, stats)
    assert f.shape == (S.N_FEAT,)
    h = f[: 8 * 64].reshape(8, 64)
    assert np.allclose(h.sum(axis=1), 8.0)
    assert h[0, 0] > 0 and h[7, 63] > 0  # the window spans the visited energies exactly
    lo, hi = 30, 154
    assert np.isclose(f[-1], np.log(2 * (hi - lo) + 2.0) / 10)
    assert np.isclose(f[-4], np.log(400) / 10) and f[-2] == 1.0
