"""Observables of --log (the writer of src/main.cc:490-530) on the GPU.

  * Parity: the engine's per-slot sums (samples, E, E^2, spins set, its
    square, |magnetisation|) equal sums taken here over the oracle's
    trajectory, step by step, on the same seed -- exactly (integers).
  * Physics: the 2D ferromagnet L = 20 (lattice_2d_20.gset) at nine
    temperatures of the exact finite-lattice tables data/lattice_2d_20_{u,c}.csv
    and at the 19 constant-acceptance temperatures of include/temps.hh:9-28
    (the exact tables interpolated, their step being 0.01): u = <E>/n within
    2e-3 and C = beta^2 var(E) / n within 3 %.  The reference ships tables
    for those 19 temperatures too (data/lattice_2d_20_cae_{u,c}.csv, which
    its run data/results.csv matches to 1.1e-3 in u); above T = 2.1 they
    depart from the exact L = 20 tables (at T = 2.6903 u = -1.0050 against
    -0.9707), so they describe some other system and are not used as the
    check.
"""
import csv
import os

import numpy as np
import pytest

from conftest import REPO, graph_path
from test_gpu_parity import make_pair

pytestmark = pytest.mark.gpu
DATA = os.path.join(REPO, "tests", "golden", "data")


def say(*a):
    print(*a, flush=True)


def table(name):
    with open(os.path.join(DATA, name)) as f:
        rows = list(csv.reader(f))
    return {round(float(r[0]), 4): float(r[1]) for r in rows[1:]}


def interp(tab, T):
    t = np.array(sorted(tab))
    return float(np.interp(T, t, [tab[x] for x in t]))


def derived(obs, betas, n):
    """u, C, chi and m per slot as the legacy writer computes them."""
    k = obs["samples"].astype(np.float64)
    e1 = obs["sum_e"] / k
    e2 = obs["sum_e2"] / k
    m1 = obs["sum_up"] / k / n
    m2 = obs["sum_up2"] / k / (n * n)
    return {"u": e1 / n, "C": betas ** 2 / n * (e2 - e1 ** 2), "chi": betas * n * (m2 - m1 ** 2), "m": m1,
            "abs_m": obs["sum_abs_m"] / k / n}


@pytest.mark.parametrize("kind", ["lattice_2d_10.gset", "pm_torus20"])
def test_observables_equal_sums_over_the_oracle_trajectory(ptx, oracle, kind):
    p, g = make_pair(ptx, oracle, kind)
    betas = oracle.beta_linspace(0.5, 3.0, 8, 0)
    start, steps = 4, 25
    eng = ptx.Engine(p, betas, seed=9, flags=ptx.FLAG_ICM | ptx.FLAG_OBSERVE, hist_start=start)
    assert eng.info().mode == ptx.MODE_SPLIT
    oe = oracle.Engine(g, betas, chains=2, seed=9, flags=oracle.FLAG_ICM)
    nrep = 2 * len(betas)
    want = {f: np.zeros(nrep, np.float64) for f in ("samples", "sum_e", "sum_e2", "sum_up", "sum_up2", "sum_abs_m")}
    for s in range(steps):
        oe.step(1)
        if s < start:
            continue
        e = oe.energies()
        up = oe.spins().sum(axis=1).astype(np.int64)
        s2r = oe.slot2rep()
        for slot in range(nrep):
            r = s2r[slot]
            want["samples"][slot] += 1
            want["sum_e"][slot] += e[r]
            want["sum_e2"][slot] += float(e[r]) ** 2
            want["sum_up"][slot] += up[r]
            want["sum_up2"][slot] += float(up[r]) ** 2
            want["sum_abs_m"][slot] += abs(2 * up[r] - p.n)
    eng.step(steps)
    got = eng.observables()
    for f, v in want.items():
        assert np.array_equal(got[f].astype(np.float64), v), f
    assert np.array_equal(eng.energies(), oe.energies())


def test_observables_need_the_split_kernels(ptx):
    p = ptx.Problem.read(graph_path("lattice_2d_10.gset"))
    with pytest.raises(ptx.PtxError) as ex:
        ptx.Engine(p, ptx.beta_linspace(0.5, 3.0, 4), flags=ptx.FLAG_OBSERVE, mode=ptx.MODE_COOP)
    assert ex.value.code == -3


def _ferro_run(ptx, temps, seed, burn=10000, samples=1000000):
    p = ptx.Problem.read(graph_path("lattice_2d_20.gset"))
    betas = 1.0 / np.asarray(temps, np.float64)
    eng = ptx.Engine(p, betas, seed=seed, flags=ptx.FLAG_OBSERVE, hist_start=burn, use_graph=50)
    eng.step(burn + samples)
    obs = eng.observables()
    eng.close()
    nt = len(temps)
    pooled = {f: v[:nt] + v[nt:] for f, v in obs.items()}  # both chains, slot by slot
    assert np.all(pooled["samples"] == 2 * samples)
    return derived(pooled, betas, p.n)


def test_ferromagnet_at_the_cae_temperatures(ptx):
    u_tab, c_tab = table("lattice_2d_20_u.csv"), table("lattice_2d_20_c.csv")
    u_cae = table("lattice_2d_20_cae_u.csv")
    temps = sorted(u_cae)
    d = _ferro_run(ptx, temps, seed=4)
    bad = []
    for i, T in enumerate(temps):
        ue, ce = interp(u_tab, T), interp(c_tab, T)
        du, dc = d["u"][i] - ue, d["C"][i] / ce - 1
        say(f"T {T}: u {d['u'][i]:.5f} exact {ue:.5f} (reference table {u_cae[T]:.5f}); "
            f"C {d['C'][i]:.4f} exact {ce:.4f}; m {d['m'][i]:.3f} |m| {d['abs_m'][i]:.3f}")
        bad += [(T, "u", du)] if abs(du) >= 2e-3 else []
        bad += [(T, "C", dc)] if abs(dc) >= 0.03 else []
    assert bad == []
    assert abs(u_cae[2.6903] - interp(u_tab, 2.6903)) > 0.03  # the reference's table is not L = 20's


def test_ferromagnet_against_the_exact_table(ptx):
    u_tab, c_tab = table("lattice_2d_20_u.csv"), table("lattice_2d_20_c.csv")
    temps = [1.5, 1.8, 2.0, 2.2, 2.3, 2.4, 2.6, 3.0, 3.5]
    d = _ferro_run(ptx, temps, seed=6)
    for i, T in enumerate(temps):
        say(f"T {T}: u {d['u'][i]:.5f} exact {u_tab[T]:.5f}; C {d['C'][i]:.4f} exact {c_tab[T]:.4f}; "
            f"chi {d['chi'][i]:.3f}; m {d['m'][i]:.3f}")
        assert abs(d["u"][i] - u_tab[T]) < 2e-3
        assert abs(d["C"][i] / c_tab[T] - 1) < 0.03
        if T >= 2.6:  # paramagnetic: as many spins set as not
            assert abs(d["m"][i] - 0.5) < 0.02
        if T <= 1.8:  # ordered: |magnetisation| near the finite-lattice value
            assert d["abs_m"][i] > 0.9
