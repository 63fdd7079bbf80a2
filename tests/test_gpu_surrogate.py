"""The trained ladder surrogate (pt-icm_amd/data/ladder_mlp.npz, trained on
the GPU by the project's own kernels on multihistogram ladders) on instances
it has not seen: what it gains over the method without it (a linear ladder
with the same number of slots), measured as adjacent swap rates in 2-chain
PT runs on the GPU, and the multihistogram fallback rule."""
import os
import sys

import numpy as np
import pytest

from conftest import REPO

sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
import surrogate as S  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _instances(ptx, n, seed):
    specs = S.sample_specs(n, seed)
    for sp in specs:  # keep the test short: the smaller half of the training ranges
        if sp["family"] == "torus":
            sp["L"] = min(sp["L"], 48)
        else:
            sp["n"] = min(sp["n"], 1600)
    coos = [S.build(sp) for sp in specs]
    return [ptx.Problem.from_coo(*c) for c in coos], [S.problem_stats(c) for c in coos]


def test_surrogate_beats_linear_ladder_on_unseen_instances(ptx):
    probs, stats = _instances(ptx, 32, seed=987)
    sur = S.Surrogate(ptx)
    X, bs, be = S.short_features(ptx, probs, stats, 0.5, 3.0, 500, 1000, seed=5)
    lads = sur.ladders(X, bs, be)
    assert all(np.all(np.diff(l) < 0) and l[0] == bs and l[-1] > be for l in lads)
    lin = [np.linspace(bs, be, len(l) + 1)[:-1] for l in lads]
    r_sur = S.measured_rates(ptx, probs, lads, 2000, seed=6)
    r_lin = S.measured_rates(ptx, probs, lin, 2000, seed=7)

    def band(rr):
        return float(np.mean(np.abs(np.concatenate(rr) - S.TARGET) <= S.TOL))
    med_sur = float(np.median([r.min() for r in r_sur]))
    med_lin = float(np.median([r.min() for r in r_lin]))
    say(f"pairs in band: surrogate {band(r_sur):.3f} linear {band(r_lin):.3f}; "
        f"median min rate: surrogate {med_sur:.3f} linear {med_lin:.3f}")
    assert band(r_sur) > band(r_lin) + 0.3
    assert med_sur > 0.2 and med_sur > 3 * med_lin


def test_fallback_rule(ptx):
    probs, stats = _instances(ptx, 12, seed=4321)
    sur = S.Surrogate(ptx)
    lads, ok, rates, t = S.choose_ladders(ptx, sur, probs, stats, seed=11, validate_steps=1000,
                                          long=(1000, 1000))
    X, bs, be = S.short_features(ptx, probs, stats, 0.5, 3.0, 500, 1000, seed=11)
    pred = sur.ladders(X, bs, be)
    for i, (lad, o, r) in enumerate(zip(lads, ok, rates)):
        assert o == S.in_band(r)
        if o:
            assert np.array_equal(lad, pred[i])  # accepted: the network's ladder is used
        else:
            assert not np.array_equal(lad, pred[i]) or len(lad) != len(pred[i])
    say("accepted", sum(ok), "of", len(ok), {k: round(v, 3) for k, v in t.items()})
