"""The trained ladder surrogate (pt-icm_amd/data/ladder_mlp.npz, trained on
the GPU by the project's own kernels on 20,000 multihistogram ladders) on
instances it has not seen, against the method it replaces -- the
multihistogram preprocess of get_beta (src/cae_estimator.cc:295-428) -- and
against a ladder linear in beta with the same number of slots: adjacent swap
rates measured in 2-chain PT runs on the GPU, and the cost of making each
ladder (pilot PT steps).  Also the multihistogram fallback rule."""
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


def test_surrogate_against_multihistogram_and_linear_on_unseen_instances(ptx):
    probs, stats = _instances(ptx, 32, seed=987)
    sur = S.Surrogate(ptx)
    X, bs, be = S.short_features(ptx, probs, stats, 0.5, 3.0, 500, 1000, seed=5)
    lads = sur.ladders(X, bs, be)
    sur.close()
    assert all(np.all(np.diff(l) < 0) and l[0] == bs and l[-1] > be for l in lads)
    mh, _, _ = S.multihistogram_ladders(ptx, probs, S.N_OUT, 0.5, 3.0, 5000, 2000, seed=8)
    lin = [np.linspace(bs, be, len(l) + 1)[:-1] for l in lads]
    r_sur = S.measured_rates(ptx, probs, lads, 2000, seed=6)
    r_mh = S.measured_rates(ptx, probs, mh, 2000, seed=6)
    r_lin = S.measured_rates(ptx, probs, lin, 2000, seed=7)

    def band(rr):
        return float(np.mean(np.abs(np.concatenate(rr) - S.TARGET) <= S.TOL))

    def all_in(rr):
        return float(np.mean([S.in_band(r) for r in rr]))

    def med_min(rr):
        return float(np.median([r.min() for r in rr]))
    # pilot cost per instance: 8 slots x 1,500 steps against 64 slots x 7,000 steps (replica sweeps)
    cost_sur, cost_mh = 8 * (500 + 1000), 64 * (5000 + 2000)
    say(f"pairs in band: surrogate {band(r_sur):.3f} multihistogram {band(r_mh):.3f} linear {band(r_lin):.3f}; "
        f"instances all in band: {all_in(r_sur):.3f} {all_in(r_mh):.3f} {all_in(r_lin):.3f}; "
        f"median min rate: {med_min(r_sur):.3f} {med_min(r_mh):.3f} {med_min(r_lin):.3f}; "
        f"pilot sweeps per replica slot: {cost_sur} vs {cost_mh}")
    # as good as the ladder it imitates, at 1/37 of its pilot cost, and far better than even spacing
    assert band(r_sur) >= band(r_mh) - 0.08
    assert all_in(r_sur) >= all_in(r_mh)
    assert med_min(r_sur) >= med_min(r_mh) - 0.05
    assert cost_mh >= 30 * cost_sur
    assert band(r_sur) > band(r_lin) + 0.3 and med_min(r_sur) > 3 * med_min(r_lin)


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
