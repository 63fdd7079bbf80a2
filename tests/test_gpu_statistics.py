"""Statistical agreement of the GPU engine with the reference's own data and
with the multihistogram machinery (the bar against the unmodified reference,
whose random streams differ): the 2D ferromagnet's energy per spin against
the tabulated u(T) the reference validates with (data/lattice_2d_20_u.csv),
and measured adjacent swap rates against the rates the multihistogram DOS
predicts for the same ladder."""
import csv
import os

import numpy as np
import pytest

from conftest import REPO, graph_path

pytestmark = pytest.mark.gpu
GOLD = os.path.join(REPO, "tests", "golden")


def test_ferromagnet_energy_per_spin_matches_reference_table(ptx):
    tab = {round(float(r["T"]), 2): float(r["u"])
           for r in csv.DictReader(open(os.path.join(GOLD, "data", "lattice_2d_20_u.csv")))}
    temps = [1.5, 2.0, 2.5, 3.0, 3.5]
    p = ptx.Problem.read(graph_path("lattice_2d_20.gset"))
    betas = 1.0 / np.array(temps)
    eng = ptx.Engine(p, betas, seed=11)
    eng.step(500)
    acc = np.zeros(2 * len(temps))
    k = 0
    for _ in range(200):
        eng.step(10)
        acc += eng.energies()[eng.slot2rep()]
        k += 1
    u = acc / k / p.n
    for i, T in enumerate(temps):
        for c in range(2):
            got = u[c * len(temps) + i]
            print(T, c, round(got, 4), tab[T], flush=True)
            assert abs(got - tab[T]) < 0.015, (T, got, tab[T])


def test_measured_swap_rates_match_multihistogram_prediction(ptx):
    """A 16-slot pilot's histograms give a DOS; the swap rates that DOS
    predicts between adjacent slots of a ladder agree with the rates a PT
    run on that ladder measures."""
    p = ptx.Problem.read(graph_path("lattice_2d_10.gset"))
    hb, bs, be = ptx.pilot_betas(0.8, 3.0, 16)
    pil = ptx.Engine(p, hb, seed=5, chains=1, flags=ptx.FLAG_PILOT | ptx.FLAG_HISTOGRAM, hist_start=2000)
    pil.step(22000)
    counts, e_base = pil.histograms()
    lad = np.sort(ptx.beta_linspace(0.8, 3.0, 10, False))[::-1].copy()
    pred = ptx.ladder_predict_rates(counts, e_base, hb, lad, mode=ptx.LADDER_NEWTON)
    eng = ptx.Engine(p, lad, seed=6)
    eng.step(20000)
    a, s, _ = eng.swap_stats()
    nt = len(lad)
    att = a[: nt - 1].astype(float) + a[nt: 2 * nt - 1]
    suc = s[: nt - 1].astype(float) + s[nt: 2 * nt - 1]
    meas = suc / att
    print(np.round(pred, 3), np.round(meas, 3), flush=True)
    assert np.all(np.abs(meas - pred) < 0.02), (pred, meas)


def test_ea_l6_ground_state_hit_rates_follow_the_reference_curve(ptx):
    """3D +-J EA L = 6 with the reference's settings (Nt = 10, beta 0.1..1.6,
    ICM): the fraction of runs whose best energy is the ground state after
    20, 54, 148, 403, 1096 and 2980 PT steps, against the reference's
    measured curve (data/pt_icm_ea_l6.csv: 0.12, 0.34, 0.78, 0.92, 0.96,
    1.00 for Nt = 10).  Its instances are not available, so this takes 10
    fresh instances x 20 runs (ground states from 20 x 20,000-step runs
    each).  Measured here: 0.375, 0.685, 0.885, 0.92, 0.95, 0.99 -- the same
    plateau from 403 steps on, reached sooner (colour-class sweeps and the
    cluster moves on the four coldest slots; the reference's early points
    rest on 50 runs of other instances).  Asserted: never more than 0.1
    below the reference, and within 0.1 of it from 403 steps on."""
    import sys
    sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
    import instances
    ref = {20: 0.12, 54: 0.34, 148: 0.78, 403: 0.92, 1096: 0.96, 2980: 1.00}
    n_inst, runs = 10, 20
    coos = [instances.ea_3d(6, seed=500 + k) for k in range(n_inst)]
    probs = [ptx.Problem.from_coo(*c) for c in coos for _ in range(runs)]
    lad = ptx.beta_linspace(0.1, 1.6, 10, True)
    long = ptx.Engine(probs, [lad] * len(probs), seed=77, flags=ptx.FLAG_ICM)
    long.step(20000)
    gs = long.min_energies().reshape(n_inst, runs).min(axis=1)
    eng = ptx.Engine(probs, [lad] * len(probs), seed=78, flags=ptx.FLAG_ICM)
    eng.step(max(ref))
    first = np.full(len(probs), np.iinfo(np.int64).max)
    for i in range(len(probs)):
        be, bs, _ = eng.best(i, with_state=False)
        hit = be == gs[i // runs]
        if hit.any():
            first[i] = bs[hit].min()
    for steps, want in ref.items():
        got = float(np.mean(first < steps))
        print(steps, round(got, 3), want, flush=True)
        assert got >= want - 0.1, (steps, got, want)
        if steps >= 403:
            assert abs(got - want) <= 0.1, (steps, got, want)
