"""The benchmark configurations on the GPU, at their real shapes.

  * A batch shaped like the mixed batched + surrogate workload: 2D +-J tori
    L = 100 (10,000 spins, cluster moves on bit sets with rows of four
    words) and Erdos-Renyi +-1 graphs G(5000, 0.004) (adjacency read through
    L2, BFS cluster moves), 64 slots x 2 chains, ICM on the cold slots.  Split
    kernels with four waves per cluster move (the autotuner's choice at
    scale), and a split engine that switches to the cooperative kernel once
    retirement leaves one instance active: bit-exact with the oracle.
  * The 4096-instance 3D EA batch (L in {6, 8, 10, 12}, 32 betas x 2 chains,
    ICM) in split mode with XCD-ordered work lists: tracked energies equal
    recomputed ones, permutations stay within their chains, every pair
    attempt is counted, and instances spread over the work lists (first,
    last, and the XCD boundaries) are bit-exact with the oracle.
  * G022 with the 64-slot pilot of --preprocess (5000 burn-in + 2000 logged
    steps): per-slot energy histograms equal the oracle's logs bit for bit,
    and the ladder equals the oracle's.
  * The pilot on a 4x4 +-J torus against the exact Boltzmann distribution of
    each slot (2^16 states enumerated).

Reference: src/solver.cc:67-127 (the step), src/cae_estimator.cc:76-151
(gather_samples), :330-428 (get_beta)."""
import os
import sys

import numpy as np
import pytest

from conftest import REPO, graph_path
from test_gpu_parity import compare
from test_ladder_pins import boltzmann, enumerate_dos, torus_couplings

sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
import instances  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _mixed_pairs(ptx, oracle, k_torus=2, k_er=2):
    pairs = []
    for i in range(k_torus + k_er):
        if i % 2 == 0 and i // 2 < k_torus:
            n, u, v, w = instances.torus_2d(100, seed=500 + i)
        else:
            n, u, v, w = instances.erdos_renyi(5000, 0.004, seed=500 + i, pm=True)
        pairs.append((ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)))
    return pairs


@pytest.mark.parametrize("shape", ["split_icm4", "split_small_batch_coop"])
def test_mixed_batch_at_scale_shapes_match_oracle(ptx, oracle, shape):
    pairs = _mixed_pairs(ptx, oracle)
    lad = np.sort(oracle.beta_linspace(0.5, 3.0, 64, 0))  # ascending: the slot rule gives the cold slots ICM
    kw = {"mode": ptx.MODE_SPLIT, "icm_waves": 4} if shape == "split_icm4" else \
        {"mode": ptx.MODE_SPLIT, "small_batch_coop": 1}
    say(shape, [p.n for p, _ in pairs])
    eng = ptx.Engine([p for p, _ in pairs], [lad] * len(pairs), seed=31, flags=ptx.FLAG_ICM, **kw)
    oes = [oracle.Engine(g, lad, chains=2, seed=31, instance=i, flags=oracle.FLAG_ICM)
           for i, (_, g) in enumerate(pairs)]
    nc = [eng.instance_info(i).nc for i in range(len(pairs))]
    assert nc == [oracle.icm_count(lad, g.char_scale) for _, g in pairs] and all(c > 0 for c in nc)
    assert pairs[0][0].info.lattice_dim == 2 and pairs[1][0].info.lattice_dim == 0
    for step in (1, 2, 5):
        eng.step(step)
        for oe in oes:
            oe.step(step)
        for i, oe in enumerate(oes):
            compare(eng, oe, inst=i, what=f"{shape} inst {i} after +{step}")
        say("ok after", step, "more steps; mode", eng.info().mode)
    if shape == "split_small_batch_coop":
        # the stragglers of a batch: one instance active at a time fits one
        # cooperative launch, and the engine switches to that kernel
        assert eng.info().coop_capacity >= 128
        for keep in (0, 1):
            active = np.zeros(len(pairs), bool)
            active[keep] = True
            eng.set_active(active)
            eng.step(3)
            oes[keep].step(3)
            assert eng.info().mode == ptx.MODE_COOP
            for i, oe in enumerate(oes):
                compare(eng, oe, inst=i, what=f"only instance {keep} active, inst {i}")
            say("cooperative steps ok for instance", keep)
    eng.close()


def test_ea_batch_of_4096_in_split_mode(ptx, oracle):
    Ls, coos = instances.ea_batch(4096, 0)
    probs = [ptx.Problem.from_coo(*c) for c in coos]
    lad = oracle.beta_linspace(0.2, 3.0, 32, 1)
    eng = ptx.Engine(probs, [lad] * len(probs), seed=0, flags=ptx.FLAG_ICM, mode=ptx.MODE_SPLIT, waves_per_group=1)
    steps = 12
    eng.step(steps)
    eng.sync()
    info = eng.info()
    assert info.mode == ptx.MODE_SPLIT and info.n_instances == 4096
    say("4096 instances stepped", steps)
    for i in range(4096):
        s2r = eng.slot2rep(i)
        assert sorted(s2r[:32].tolist()) == list(range(32)) and sorted(s2r[32:].tolist()) == list(range(32, 64)), i
        a, s, nsw = eng.swap_stats(i)
        assert a.sum() == steps * 2 * 31 and s.sum() == nsw, i
    rng = np.random.default_rng(1)
    for i in rng.choice(4096, 48, replace=False):
        sp = eng.spins(int(i))
        assert np.array_equal(eng.energies(int(i)), [probs[int(i)].energy(x) for x in sp]), i
    # instances at the ends of the size order and around XCD boundaries of the work lists
    for i in (0, 1, 7, 8, 9, 2047, 2048, 4095):
        g = oracle.Graph.from_coo(*coos[i])
        oe = oracle.Engine(g, lad, chains=2, seed=0, instance=i, flags=oracle.FLAG_ICM)
        oe.step(steps)
        compare(eng, oe, inst=i, what=f"EA instance {i} (L={Ls[i]})")
    say("bit-exact sample ok")
    eng.close()


def test_g022_preprocess_pilot_matches_oracle(ptx, oracle):
    p, g = ptx.Problem.read(graph_path("G022")), oracle.Graph.read(graph_path("G022"))
    nt, burn, ns = 64, 5000, 2000
    hb, bs, be = ptx.pilot_betas(0.6, 2.8, nt)
    eng = ptx.Engine(p, hb, seed=0, chains=1, flags=ptx.FLAG_PILOT | ptx.FLAG_HISTOGRAM, hist_start=burn)
    eng.step(burn + ns)
    counts, e_base = eng.histograms()
    say("GPU pilot done; mode", eng.info().mode)
    eng.close()
    logs = oracle.gather_samples(g, hb, 0, burn, ns)
    say("oracle pilot done")
    assert counts.sum() == ns * nt
    for j in range(nt):
        want = np.bincount((logs[:, j] - e_base) // 2, minlength=counts.shape[1])
        assert np.array_equal(counts[j], want), j
    fast = ptx.ladder_from_histograms(counts, e_base, hb, bs, be, mode=ptx.LADDER_FAST)
    olad, _, _ = oracle.get_beta_from_logs(logs, hb, 2.0, 0.1, 0.4, bs, be)
    say("ladders", len(fast), len(olad))
    assert len(fast) == len(olad) and np.allclose(fast, olad, rtol=1e-12, atol=0)


@pytest.mark.parametrize("mode", [1, 3])
def test_gpu_pilot_slots_follow_exact_boltzmann(ptx, mode):
    b, w = torus_couplings(4, True, 3)
    E, lg = enumerate_dos(16, b, w)
    p = ptx.Problem.from_coo(16, b[:, 0].astype(np.int32), b[:, 1].astype(np.int32), w.astype(np.float64))
    hb, _, _ = ptx.pilot_betas(0.6, 3.0, 8)
    burn, ns = 2000, 200000
    eng = ptx.Engine(p, hb, seed=5, chains=1, flags=ptx.FLAG_PILOT | ptx.FLAG_HISTOGRAM, hist_start=burn, mode=mode)
    eng.step(burn + ns)
    counts, e_base = eng.histograms()
    eng.close()
    for j, beta in enumerate(hb):
        c = counts[j].astype(np.float64)
        assert c.sum() == ns
        emp = np.zeros(len(E))
        for k, e in enumerate(E):
            emp[k] = c[(e - e_base) // 2]
        assert emp.sum() == ns, "counts outside the exact spectrum"
        pe = boltzmann(E, lg, beta)
        tvd = 0.5 * np.abs(emp / ns - pe).sum()
        mean_emp, mean_exact = float((emp * E).sum() / ns), float((pe * E).sum())
        say(j, round(float(beta), 3), "tvd", round(tvd, 4), "mean", round(mean_emp, 3), round(mean_exact, 3))
        assert tvd < 0.02
        assert abs(mean_emp - mean_exact) < 0.1
