"""Pins of the oracle's ladder search and pilot run that do not lean on the
restatement itself.

The ladder search (get_beta, reference src/cae_estimator.cc:360-428) is fed
energy histograms of systems whose density of states is known exactly: an
ideal paramagnet of N spins in a unit field (g(E) = C(N, (N + E)/2)) and 4x4
periodic lattices (2^16 states enumerated here).  Whatever the step schedule
does, the loop's text fixes what its output must satisfy, and these tests
check that against acceptance rates computed here from the exact g(E) by a
direct double sum:

  * ladder[0] = beta_start, strictly decreasing, every entry above beta_end;
  * a kept pair's rate lies in [r - 0.05, r]: the loop backs off whenever
    r - rate > 0.05 (:397-402) and keeps a temperature only once rate <= r
    or beta <= beta_end (:413), and a temperature at or below beta_end is
    dropped (:414-416);
  * the ladder ends early only when beta_end is within reach: the trial that
    crossed beta_end was not backed off, so its rate was >= r - 0.05, and
    the rate against beta_end itself is larger still;
  * every temperature step 1/beta_i - 1/beta_{i-1} is step_size times a sum
    of powers of two (the step only halves or doubles, :400, :405);
  * two cases are worked out from the closed form: on a 256-spin paramagnet
    at T0 = 0.8 the first trial T0 + 0.1 has rate 0.379, inside the band,
    so beta_1 = 1/(1/beta_0 + 0.1); on a 576-spin paramagnet at T0 = 0.5 the
    first trial has rate 0.078 (backed off, step 0.05) and T0 + 0.05 has
    0.373, so beta_1 = 1/(1/beta_0 + 0.05).

Planted mistakes -- the back-off test with its sign flipped, the beta pair
of the acceptance swapped, the temperature stepped the wrong way -- are
compiled into copies of the oracle here and must fail these checks.

The pilot run (gather_samples, cae_estimator.cc:76-151) is pinned by its
per-slot energy distributions: on a 4x4 +-J torus each slot's logged
energies follow the exact Boltzmann distribution at the slot's beta
(chi-square on thinned samples), and a copy of the oracle with the exchange
rule's sign flipped fails the same test.

The product's ladder modes (exact, fast, Newton) must pass the same checks.
"""
import ctypes as C
import math
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest
from scipy.special import gammaln
from scipy.stats import chi2

from conftest import REPO

R, STEP = 0.4, 0.1
TOL = 2e-3  # exact-DOS rates against the algorithm's own (WHAM of rounded counts)
ORACLE_C = os.path.join(REPO, "oracle", "oracle.c")


# ------------------------------------------------------------ exact systems --
def paramagnet(N):
    k = np.arange(N + 1)
    return (-N + 2 * k).astype(np.int64), gammaln(N + 1) - gammaln(k + 1) - gammaln(N - k + 1)


def torus_bonds(L):
    idx = np.arange(L * L).reshape(L, L)
    return np.concatenate([np.stack([idx.ravel(), np.roll(idx, -1, 1).ravel()], 1),
                           np.stack([idx.ravel(), np.roll(idx, -1, 0).ravel()], 1)])


def torus_couplings(L, pm, seed):
    b = torus_bonds(L)
    w = np.where(np.random.default_rng(seed).random(len(b)) < 0.5, -1, 1) if pm else -np.ones(len(b), np.int64)
    return b, w


def enumerate_dos(n, bonds, w):
    """E = sum over edges of w s_i s_j (the reference's energy, problem.cc:121-142), all 2^n states."""
    x = np.arange(2 ** n, dtype=np.int64)
    s = ((x[:, None] >> np.arange(n)) & 1) * 2 - 1
    e = (w[None, :] * s[:, bonds[:, 0]] * s[:, bonds[:, 1]]).sum(1)
    vals, cnt = np.unique(e, return_counts=True)
    return vals.astype(np.int64), np.log(cnt.astype(np.float64))


def boltzmann(E, logg, beta):
    a = logg - beta * E
    a = a - a.max()
    p = np.exp(a)
    return p / p.sum()


def exact_rate(E, logg, b1, b2):
    """Swap acceptance sum_ij P1_i P2_j min(1, exp((b1 - b2)(E_i - E_j))) by a direct double sum."""
    p1, p2 = boltzmann(E, logg, b1), boltzmann(E, logg, b2)
    x = (b1 - b2) * (E[:, None] - E[None, :]).astype(np.float64)
    return float(p1 @ np.exp(np.minimum(x, 0.0)) @ p2)


def exact_counts(p, n):
    """n samples distributed over the energies as n p, rounded by largest remainder (sum exactly n)."""
    raw = p * n
    c = np.floor(raw).astype(np.int64)
    rest = n - c.sum()
    c[np.argsort(-(raw - c), kind="stable")[:rest]] += 1
    return c


def exact_histograms(E, logg, hb, n):
    """Per pilot slot: counts of every energy of E (exact Boltzmann proportions)."""
    return np.stack([exact_counts(boltzmann(E, logg, b), n) for b in hb])


def logs_from_counts(E, counts):
    """Energy log [n][nt] whose per-slot histograms are the given counts."""
    return np.stack([np.repeat(E, c) for c in counts], 1).astype(np.int32)


def product_counts(E, counts):
    """The engine's histogram layout: energy = e_base + 2 * bin."""
    e_base = int(E.min())
    nb = int((E.max() - e_base) // 2) + 1
    out = np.zeros((counts.shape[0], nb), np.uint32)
    out[:, (E - e_base) // 2] = counts
    return out, e_base


# --------------------------------------------------------------- oracle ----
def build_oracle_copy(tmp_path, name, edits=()):
    """A copy of the oracle with textual edits (planted mistakes), built here."""
    src = open(ORACLE_C).read()
    for old, new in edits:
        assert src.count(old) == 1, old
        src = src.replace(old, new)
    d = tmp_path / name
    d.mkdir()
    shutil.copy(os.path.join(REPO, "oracle", "oracle.h"), d / "oracle.h")
    (d / "oracle.c").write_text(src)
    so = d / "liboracle.so"
    subprocess.run(["gcc", "-std=c11", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-o",
                    str(so), str(d / "oracle.c"), "-lm"], check=True)
    return _load_copy(str(so))


def _load_copy(so):
    L = C.CDLL(so)
    P = C.POINTER
    L.or_get_beta_from_logs.restype = C.c_int32
    L.or_get_beta_from_logs.argtypes = [P(C.c_int32), C.c_int64, C.c_int32, P(C.c_double), C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, P(C.c_double), P(C.c_double), C.c_int64,
                                        P(C.c_int64), P(C.c_double)]
    L.or_acceptance_rate.restype = C.c_double
    L.or_acceptance_rate.argtypes = [P(C.c_double)] * 3 + [C.c_int64] + [C.c_double] * 4
    L.or_partition_func.restype = C.c_double
    L.or_partition_func.argtypes = [P(C.c_double), C.c_int64, C.c_double, C.c_double, C.c_double]
    L.or_graph_from_coo.restype = C.c_void_p
    L.or_graph_from_coo.argtypes = [C.c_int32, C.c_int64, P(C.c_int32), P(C.c_int32), P(C.c_double),
                                    P(C.c_double), C.c_int, C.c_char_p, C.c_int]
    L.or_graph_free.argtypes = [C.c_void_p]
    L.or_gather_samples.argtypes = [C.c_void_p, P(C.c_double), C.c_int32, C.c_uint64, C.c_uint32, C.c_int64,
                                    C.c_int64, P(C.c_int32), C.c_uint32]
    return L


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def oracle_ladder(L, elog, hb, bs, be, nt=None):
    """or_get_beta_from_logs; returns the ladder and the loop's own rates of its adjacent pairs
    (the same DOS and the same acceptance routine)."""
    elog = np.ascontiguousarray(elog, np.int32)
    ns = elog.shape[0]
    nt = nt or elog.shape[1]
    hb = np.ascontiguousarray(hb, np.float64)
    out = np.zeros(max(nt, elog.shape[1]), np.float64)
    cap = 1 << 16
    logdos = np.zeros(cap, np.float64)
    nb = C.c_int64(0)
    logz = np.zeros(elog.shape[1], np.float64)
    k = L.or_get_beta_from_logs(_p(elog, C.c_int32), ns, elog.shape[1], _p(hb, C.c_double), 2.0, STEP, R, bs, be,
                                _p(out, C.c_double), _p(logdos, C.c_double), cap, C.byref(nb), _p(logz, C.c_double))
    betas = out[:k].copy()
    ld = np.ascontiguousarray(logdos[: nb.value])
    emin = float(elog.min())
    e = emin + 2.0 * np.arange(len(ld))

    def pe(b):
        z = L.or_partition_func(_p(ld, C.c_double), len(ld), b, emin, 2.0)
        return np.ascontiguousarray(ld - e * b - z)

    own = []
    for b1, b2 in zip(betas[:-1], betas[1:]):
        p1, p2 = pe(b1), pe(b2)
        own.append(L.or_acceptance_rate(_p(ld, C.c_double), _p(p1, C.c_double), _p(p2, C.c_double), len(ld), b1, b2,
                                        2.0, emin))
    return betas, np.array(own)


def ladder_failures(betas, own, E, logg, bs, be, nt):
    """The checks listed in the module docstring; returns the list of violated ones."""
    bad = []
    if len(betas) < 1 or betas[0] != bs:
        bad.append("ladder[0] != beta_start")
    if not np.all(np.diff(betas) < 0):
        bad.append("not strictly decreasing")
    if len(betas) and not np.all(betas > be):
        bad.append("entry at or below beta_end")
    if len(own) and not (np.all(own >= R - 0.05 - 1e-9) and np.all(own <= R + 1e-9)):
        bad.append(f"own rates outside the band: {own}")
    ex = np.array([exact_rate(E, logg, a, b) for a, b in zip(betas[:-1], betas[1:])])
    if len(ex) and not (np.all(ex >= R - 0.05 - TOL) and np.all(ex <= R + TOL)):
        bad.append(f"exact rates outside the band: {ex}")
    if 1 <= len(betas) < nt and exact_rate(E, logg, betas[-1], be) < R - 0.05 - TOL:
        bad.append("stopped before nt with beta_end out of reach")
    for a, b in zip(betas[:-1], betas[1:]):
        q = (1.0 / b - 1.0 / a) / STEP * 65536.0
        if q <= 0 or abs(q - round(q)) > 1e-4 * max(1.0, q / 65536.0):
            bad.append(f"temperature step {1.0 / b - 1.0 / a} is not step_size times a dyadic number")
            break
    return bad


# systems: (name, E, logg, T0, T1, pilot slots, ladder slots, samples per slot)
def _systems():
    b, w = torus_couplings(4, True, 3)
    E_pm, lg_pm = enumerate_dos(16, b, w)
    b, w = torus_couplings(4, False, 0)
    E_f, lg_f = enumerate_dos(16, b, w)
    E64, lg64 = paramagnet(64)
    E256, lg256 = paramagnet(256)
    E576, lg576 = paramagnet(576)
    return {
        "torus_pmJ_4x4": (E_pm, lg_pm, 0.5, 3.0, 8, 16, 400000),
        "torus_ferro_4x4": (E_f, lg_f, 1.0, 4.0, 8, 16, 400000),
        "paramagnet_64": (E64, lg64, 0.5, 5.0, 8, 4, 400000),
        "paramagnet_256": (E256, lg256, 0.8, 1.6, 8, 8, 400000),
        "paramagnet_576": (E576, lg576, 0.5, 1.0, 8, 8, 400000),
    }


SYSTEMS = _systems()


def _inputs(oracle, name):
    E, lg, T0, T1, nh, nt, ns = SYSTEMS[name]
    hb, bs, be = oracle.hist_betas(T0, T1, nh, 0)
    counts = exact_histograms(E, lg, hb, ns)
    return E, lg, hb, bs, be, nt, counts


@pytest.fixture(scope="module")
def base_oracle(tmp_path_factory):
    return build_oracle_copy(tmp_path_factory.mktemp("oracle"), "base")


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_oracle_ladder_invariants_on_exact_dos(oracle, base_oracle, name):
    E, lg, hb, bs, be, nt, counts = _inputs(oracle, name)
    elog = logs_from_counts(E, counts)
    betas, own = oracle_ladder(base_oracle, elog, hb, bs, be)
    assert ladder_failures(betas, own, E, lg, bs, be, elog.shape[1]) == []
    assert len(betas) >= 2


def test_oracle_ladder_exercises_both_endings(oracle, base_oracle):
    """Some system stops early (beta_end reached) and some fills every slot."""
    ends = set()
    for name in SYSTEMS:
        E, lg, hb, bs, be, nt, counts = _inputs(oracle, name)
        betas, _ = oracle_ladder(base_oracle, logs_from_counts(E, counts), hb, bs, be)
        ends.add(len(betas) < len(hb))
    assert ends == {True, False}


def test_closed_form_first_trial_kept(oracle, base_oracle):
    """256-spin paramagnet, T0 = 0.8: the exact rate of (T0, T0 + 0.1) is
    0.379, inside [0.35, 0.4], so the first trial is kept."""
    E, lg, hb, bs, be, nt, counts = _inputs(oracle, "paramagnet_256")
    assert bs == 1.25
    r01 = exact_rate(E, lg, bs, 1.0 / (1.0 / bs + 0.1))
    assert 0.37 < r01 < 0.39
    betas, _ = oracle_ladder(base_oracle, logs_from_counts(E, counts), hb, bs, be)
    assert betas[1] == 1.0 / (1.0 / bs + 0.1)


def test_closed_form_back_off(oracle, base_oracle):
    """576-spin paramagnet, T0 = 0.5: the first trial T0 + 0.1 has exact rate
    0.078, more than 0.05 below r, so the loop backs off and halves the step;
    the next trial T0 + 0.05 has rate 0.373, inside the band, and is kept."""
    E, lg, hb, bs, be, nt, counts = _inputs(oracle, "paramagnet_576")
    assert bs == 2.0
    assert exact_rate(E, lg, bs, 1.0 / (1.0 / bs + 0.1)) < 0.1
    assert 0.36 < exact_rate(E, lg, bs, 1.0 / (1.0 / bs + 0.05)) < 0.39
    betas, _ = oracle_ladder(base_oracle, logs_from_counts(E, counts), hb, bs, be)
    assert betas[1] == 1.0 / (1.0 / bs + 0.1 / 2)


MUTANTS = {
    "back-off sign flipped": [("if ((r - r_curr) > 0.05) {", "if ((r_curr - r) > 0.05) {")],
    "acceptance beta pair swapped": [("double p12 = (beta_1 - beta_2) * (e1 - e2);",
                                      "double p12 = (beta_2 - beta_1) * (e1 - e2);")],
    "temperature stepped down": [("beta_curr = 1 / (1 / beta_curr + step);",
                                  "beta_curr = 1 / (1 / beta_curr - step);")],
}


def mutant_failures(so_path, name):
    """The checks on one system with the oracle copy at so_path (run in a child
    process: a planted mistake may keep the search from terminating)."""
    import oracle_lib
    L = _load_copy(so_path)
    E, lg, hb, bs, be, nt, counts = _inputs(oracle_lib, name)
    betas, own = oracle_ladder(L, logs_from_counts(E, counts), hb, bs, be)
    bad = ladder_failures(betas, own, E, lg, bs, be, len(hb))
    if name == "paramagnet_256" and (len(betas) < 2 or betas[1] != 1.0 / (1.0 / bs + 0.1)):
        bad.append("closed-form first trial")
    if name == "paramagnet_576" and (len(betas) < 2 or betas[1] != 1.0 / (1.0 / bs + 0.05)):
        bad.append("closed-form back-off")
    return bad


@pytest.mark.parametrize("mutant", list(MUTANTS))
def test_planted_ladder_mistakes_fail(tmp_path, mutant):
    build_oracle_copy(tmp_path, "m", MUTANTS[mutant])
    so = str(tmp_path / "m" / "liboracle.so")
    failed = []
    for name in SYSTEMS:
        code = ("import sys; sys.path.insert(0, %r); import test_ladder_pins as T; "
                "print(repr(T.mutant_failures(%r, %r)))" % (os.path.join(REPO, "tests"), so, name))
        try:  # the unmodified search takes well under a second on each of these systems
            r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=15)
        except subprocess.TimeoutExpired:
            failed.append((name, "search did not terminate within 15 s"))
            break
        if r.returncode != 0:
            failed.append((name, r.stderr.strip().splitlines()[-1:] or ["child failed"]))
        elif r.stdout.strip() != "[]":
            failed.append((name, r.stdout.strip()))
        if failed:
            break
    assert failed, f"planted mistake '{mutant}' passed every check"


@pytest.mark.parametrize("mode", ["exact", "fast", "newton"])
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_product_ladder_invariants_on_exact_dos(ptx, oracle, name, mode):
    E, lg, hb, bs, be, nt, counts = _inputs(oracle, name)
    pc, e_base = product_counts(E, counts)
    m = {"exact": ptx.LADDER_EXACT, "fast": ptx.LADDER_FAST, "newton": ptx.LADDER_NEWTON}[mode]
    betas = ptx.ladder_from_histograms(pc, e_base, hb, bs, be, mode=m)
    own = ptx.ladder_predict_rates(pc, e_base, hb, betas, mode=m) if len(betas) > 1 else np.zeros(0)
    # fast and Newton modes compute the same sums in another order (own rates within 1e-9 of the exact ones)
    assert ladder_failures(betas, own, E, lg, bs, be, len(hb)) == []
    if name == "paramagnet_256":
        assert betas[1] == 1.0 / (1.0 / bs + 0.1)
    if name == "paramagnet_576":
        assert betas[1] == 1.0 / (1.0 / bs + 0.05)


# ------------------------------------------------------------------ pilot ---
def _pilot_graph(L):
    b, w = torus_couplings(4, True, 3)
    u = np.ascontiguousarray(b[:, 0], np.int32)
    v = np.ascontiguousarray(b[:, 1], np.int32)
    ww = np.ascontiguousarray(w, np.float64)
    err = C.create_string_buffer(256)
    g = L.or_graph_from_coo(16, len(u), _p(u, C.c_int32), _p(v, C.c_int32), _p(ww, C.c_double), None, 0, err, 256)
    assert g, err.value
    return g, b, w


def pilot_chi2_pvalues(L, hb, seed=11, burn=2000, ns=60000, thin=6):
    g, b, w = _pilot_graph(L)
    E, lg = enumerate_dos(16, b, w)
    hb = np.ascontiguousarray(hb, np.float64)
    elog = np.zeros((ns, len(hb)), np.int32)
    L.or_gather_samples(g, _p(hb, C.c_double), len(hb), seed, 0, burn, ns, _p(elog, C.c_int32), 0)
    L.or_graph_free(g)
    out = []
    for j, beta in enumerate(hb):
        x = elog[::thin, j]
        p = boltzmann(E, lg, beta)
        obs = np.array([(x == e).sum() for e in E], np.float64)
        assert obs.sum() == len(x), "a logged energy outside the exact spectrum"
        exp_ = p * len(x)
        # merge energies (in order) until each group expects at least 5 samples
        groups_o, groups_e, co, ce = [], [], 0.0, 0.0
        for o, e in zip(obs, exp_):
            co, ce = co + o, ce + e
            if ce >= 5:
                groups_o.append(co)
                groups_e.append(ce)
                co, ce = 0.0, 0.0
        if ce > 0:
            groups_o[-1] += co
            groups_e[-1] += ce
        go, ge = np.array(groups_o), np.array(groups_e)
        stat = float(((go - ge) ** 2 / ge).sum())
        out.append(float(chi2.sf(stat, max(len(go) - 1, 1))))
    return np.array(out)


def test_pilot_slots_follow_exact_boltzmann(oracle, base_oracle):
    hb, _, _ = oracle.hist_betas(0.6, 3.0, 8, 0)
    pv = pilot_chi2_pvalues(base_oracle, hb)
    assert np.all(pv > 1e-3), pv


def test_planted_pilot_exchange_mistake_fails(oracle, tmp_path):
    L = build_oracle_copy(tmp_path, "mx", [("double dB = e->beta[k] - e->beta[k + 1];",
                                            "double dB = e->beta[k + 1] - e->beta[k];")])
    hb, _, _ = oracle.hist_betas(0.6, 3.0, 8, 0)
    pv = pilot_chi2_pvalues(L, hb)
    assert np.any(pv < 1e-6), pv
