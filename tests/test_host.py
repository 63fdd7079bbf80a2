"""CPU tests of the product's host side through the C ABI (no GPU calls):
exported symbols, reader/energy/colouring/ladder parity with the oracle, and
error behaviour at the boundary."""
import math
import os
import re
import sys

import numpy as np
import pytest

from conftest import REPO, graph_path

sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
import instances  # noqa: E402

GRAPHS = ["lattice_2d_5.gset", "lattice_2d_10.gset", "lattice_2d_20.gset", "G022", "G039"]


def test_every_declared_symbol_is_exported(ptx):
    hdr = open(os.path.join(REPO, "include", "ptx.h")).read()
    declared = set(re.findall(r"^(?:ptx_status|void|int32_t|const char\*)\s+(ptx_[a-z0-9_]+)\s*\(", hdr, re.M))
    L = ptx.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert declared <= set(ptx.exported_symbols()) | {"ptx_engine_config_default"}


@pytest.mark.parametrize("name", GRAPHS)
def test_reader_matches_oracle(ptx, oracle, name):
    p = ptx.Problem.read(graph_path(name))
    g = oracle.Graph.read(graph_path(name))
    assert p.n == g.n and p.info.nnz == g.nnz and bool(p.info.sparse) == g.sparse
    assert p.info.char_scale == g.char_scale
    for a, b in zip(p.csr(), g.csr()):
        assert np.array_equal(a, b)
    k, c = g.colour()
    p2o, cs = g.positions()
    q2o, qs = p.colouring()
    assert p.info.n_colours == k and np.array_equal(p2o, q2o) and np.array_equal(cs, qs)
    assert p.info.fmax == g.fmax()
    rng = np.random.default_rng(0)
    for _ in range(4):
        s = rng.integers(0, 2, p.n).astype(np.uint8)
        assert p.energy(s) == g.energy(s)
        assert p.maxcut(s) == g.maxcut(s)
        d, e = p.delta_e(s)
        assert e == g.energy(s) and np.array_equal(d, g.delta_e(s))


def test_coo_field_and_dense(ptx, oracle):
    rng = np.random.default_rng(4)
    n = 40
    u, v = np.triu_indices(n, 1)
    keep = rng.random(len(u)) < 0.3
    u, v = u[keep].astype(np.int32), v[keep].astype(np.int32)
    w = rng.integers(-3, 4, len(u)).astype(np.float64)
    h = rng.integers(-2, 3, n).astype(np.float64)
    for dense in (False, True):
        p = ptx.Problem.from_coo(n, u, v, w, h, dense=dense)
        g = oracle.Graph.from_coo(n, u, v, w, h, dense=dense)
        for a, b in zip(p.csr(), g.csr()):
            assert np.array_equal(a, b)
        s = rng.integers(0, 2, n).astype(np.uint8)
        assert p.energy(s) == g.energy(s)
        assert p.info.has_field == 1


def test_ladders_match_oracle(ptx, oracle):
    for T0, T1, n, inv in [(0.5, 3.0, 32, 0), (2, 0.25, 20, 0), (0.2, 3.0, 32, 1), (3.0, 0.2, 7, 1)]:
        assert np.array_equal(ptx.beta_linspace(T0, T1, n, inv), oracle.beta_linspace(T0, T1, n, inv))
        hb, bs, be = ptx.pilot_betas(T0, T1, n, inv)
        hb2, bs2, be2 = oracle.hist_betas(T0, T1, n, inv)
        assert np.array_equal(hb, hb2) and bs == bs2 and be == be2
    b = np.linspace(0.2, 3.0, 29)
    for cs in (0.1, 1.0, 2.9, 5.0):
        assert ptx.icm_count(b, cs) == oracle.icm_count(b, cs)


def test_portable_fp64_host_matches_oracle(ptx, oracle):
    L = oracle.lib()
    rng = np.random.default_rng(5)
    x = np.concatenate([-rng.random(20000) * 60, -rng.random(2000) * 745, [0.0, -1e-300, -745.0, -708.0]])
    e, _ = ptx.host_fp64(x)
    assert all(e[i] == L.or_exp(float(x[i])) for i in range(len(x)))
    u = np.concatenate([rng.integers(1, 2 ** 32, 20000) * 2.0 ** -32, [1.0, 0.5, 2 ** -32, 1 - 2 ** -32]])
    _, lg = ptx.host_fp64(u)
    assert all(lg[i] == L.or_log(float(u[i])) for i in range(len(u)))


def _pilot_logs(oracle, name, nt, ns, seed=0):
    g = oracle.Graph.read(graph_path(name))
    hb, bs, be = oracle.hist_betas(1.0, 3.0, nt, 0)
    logs = oracle.gather_samples(g, hb, seed, 200, ns)
    return g, hb, bs, be, logs


def _logs_to_engine_hist(logs, e_base):
    nt = logs.shape[1]
    nb = int((logs.max() - e_base) // 2) + 1
    h = np.zeros((nt, nb), np.uint32)
    for j in range(nt):
        b = (logs[:, j] - e_base) // 2
        np.add.at(h[j], b, 1)
    return h


def test_ladder_from_histograms_matches_oracle_exactly(ptx, oracle):
    g, hb, bs, be, logs = _pilot_logs(oracle, "lattice_2d_10.gset", 8, 3000)
    want, logdos, logz = oracle.get_beta_from_logs(logs, hb, 2.0, 0.1, 0.4, bs, be)
    e_base = -200
    counts = _logs_to_engine_hist(logs, e_base)
    got = ptx.ladder_from_histograms(counts, e_base, hb, bs, be, mode=ptx.LADDER_EXACT)
    assert np.array_equal(got, want)
    fast = ptx.ladder_from_histograms(counts, e_base, hb, bs, be, mode=ptx.LADDER_FAST)
    assert len(fast) == len(want) and np.allclose(fast, want, rtol=1e-9)


def test_fast_acceptance_equals_exact(ptx, oracle):
    g, hb, bs, be, logs = _pilot_logs(oracle, "lattice_2d_20.gset", 10, 2000, seed=3)
    e_base = -800
    counts = _logs_to_engine_hist(logs, e_base)
    lad = np.linspace(hb[0], hb[-1], 12)
    r_exact = ptx.ladder_predict_rates(counts, e_base, hb, lad, mode=ptx.LADDER_EXACT)
    r_fast = ptx.ladder_predict_rates(counts, e_base, hb, lad, mode=ptx.LADDER_FAST)
    assert np.allclose(r_exact, r_fast, rtol=1e-10, atol=1e-14)
    assert np.all((r_exact > 0) & (r_exact <= 1))


def _torus4_dos():
    """Exact density of states of the 4x4 periodic ferromagnet (w = -1, E = sum w s s)."""
    L = 4
    idx = np.arange(16).reshape(L, L)
    edges = np.concatenate([np.stack([idx.ravel(), np.roll(idx, -1, 1).ravel()], 1),
                            np.stack([idx.ravel(), np.roll(idx, -1, 0).ravel()], 1)])
    states = ((np.arange(1 << 16)[:, None] >> np.arange(16)) & 1) * 2 - 1
    E = -np.sum(states[:, edges[:, 0]] * states[:, edges[:, 1]], axis=1)
    ev, g = np.unique(E, return_counts=True)
    return ev.astype(float), g.astype(float)


def _exact_rate(ev, g, b1, b2):
    P1 = g * np.exp(-b1 * ev); P1 /= P1.sum()
    P2 = g * np.exp(-b2 * ev); P2 /= P2.sum()
    hi, lo = max(b1, b2), min(b1, b2)
    Pa, Pb = (P1, P2) if b1 >= b2 else (P2, P1)
    return float(sum(Pa[i] * Pb[j] * min(1.0, np.exp((hi - lo) * (ev[i] - ev[j])))
                     for i in range(len(ev)) for j in range(len(ev))))


def test_newton_and_iteration_recover_exact_acceptance(ptx):
    """Histograms proportional to the exact 4x4 DOS times Boltzmann factors:
    both multihistogram solvers must predict the exact swap rates; the Newton
    solve (true fixed point) at least as closely as the reference iteration."""
    ev, g = _torus4_dos()
    hb = np.array([0.9, 0.7, 0.5, 0.35, 0.2])
    n = 2_000_000
    e_base = -32
    counts = np.zeros((len(hb), 33), np.uint32)
    for k, b in enumerate(hb):
        P = g * np.exp(-b * ev); P /= P.sum()
        counts[k, ((ev - e_base) // 2).astype(int)] = np.rint(P * n).astype(np.uint32)
    lad = np.array([0.85, 0.6, 0.45, 0.3, 0.22])
    want = np.array([_exact_rate(ev, g, lad[k], lad[k + 1]) for k in range(len(lad) - 1)])
    r_newton = ptx.ladder_predict_rates(counts, e_base, hb, lad, mode=ptx.LADDER_NEWTON)
    r_iter = ptx.ladder_predict_rates(counts, e_base, hb, lad, mode=ptx.LADDER_EXACT)
    err_n = np.max(np.abs(r_newton - want))
    err_i = np.max(np.abs(r_iter - want))
    assert err_n < 1e-5, (r_newton, want)  # exact up to the rounding of the counts
    assert err_n <= err_i + 1e-4


def test_newton_ladder_close_to_reference_iteration(ptx, oracle):
    g2, hb2, bs2, be2, logs2 = _pilot_logs(oracle, "lattice_2d_10.gset", 8, 3000)
    c2 = _logs_to_engine_hist(logs2, -200)
    want = ptx.ladder_from_histograms(c2, -200, hb2, bs2, be2, mode=ptx.LADDER_EXACT)
    got = ptx.ladder_from_histograms(c2, -200, hb2, bs2, be2, mode=ptx.LADDER_NEWTON)
    assert abs(len(got) - len(want)) <= 1
    k = min(len(got), len(want))
    assert np.allclose(got[:k], want[:k], atol=0.03)


# ------------------------------------------------------------- errors -------
def test_errors_are_status_codes(ptx, tmp_path):
    with pytest.raises(ptx.PtxError) as ei:
        ptx.Problem.read(str(tmp_path / "missing.gset"))
    assert ei.value.code == -2
    bad = tmp_path / "bad.gset"
    bad.write_text("3 2\n0 1 1\nx y z\n")
    with pytest.raises(ptx.PtxError) as ei:
        ptx.Problem.read(str(bad))
    assert ei.value.code == -1
    frac = tmp_path / "frac.gset"
    frac.write_text("3 2\n0 1 0.5\n1 2 1\n")
    with pytest.raises(ptx.PtxError) as ei:
        ptx.Problem.read(str(frac))
    assert ei.value.code == -3
    trunc = tmp_path / "trunc.gset"
    trunc.write_text("3 2\n0 7 1\n")
    with pytest.raises(ptx.PtxError) as ei:
        ptx.Problem.read(str(trunc))
    assert ei.value.code == -1


def test_engine_without_device_fails_loudly(ptx):
    if ptx.lib().ptx_device_count() > 0:
        pytest.skip("a GPU is present")
    p = ptx.Problem.read(graph_path("lattice_2d_10.gset"))
    with pytest.raises(ptx.PtxError) as ei:
        ptx.Engine(p, ptx.beta_linspace(0.5, 3.0, 4))
    assert ei.value.code == -4


def _lattice_edges(dim, L, seed=0):
    idx = np.arange(L ** dim).reshape((L,) * dim)
    e = np.concatenate([np.stack([idx.ravel(), np.roll(idx, -1, axis=a).ravel()], 1) for a in range(dim)])
    w = np.where(np.random.default_rng(seed).integers(0, 2, len(e)) == 1, 1.0, -1.0)
    return L ** dim, e[:, 0].astype(np.int32), e[:, 1].astype(np.int32), w


@pytest.mark.parametrize("dim,L", [(2, 3), (2, 5), (2, 20), (2, 33), (2, 64), (3, 3), (3, 6), (3, 11)])
def test_lattice_detection(ptx, dim, L):
    """Periodic L^dim lattices numbered like their axes are recognised in any
    edge order and with zero weights; anything else is not."""
    n, u, v, w = _lattice_edges(dim, L, seed=L)
    rng = np.random.default_rng(1)
    perm = rng.permutation(len(u))
    flip = rng.integers(0, 2, len(u)) == 1
    uu = np.where(flip, v, u)[perm]
    vv = np.where(flip, u, v)[perm]
    ww = w[perm].copy()
    ww[:3] = 0.0  # zero-weight entries stay edges of a sparse instance
    p = ptx.Problem.from_coo(n, uu, vv, ww)
    assert (p.info.lattice_dim, p.info.lattice_L) == (dim, L)
    # a missing bond, a doubled bond, an extra bond, a relabelled site: no lattice
    for uu2, vv2 in [(uu[1:], vv[1:]), (np.append(uu, uu[0]), np.append(vv, vv[0])),
                     (np.append(uu, 0), np.append(vv, n // 2 + 1))]:
        q = ptx.Problem.from_coo(n, uu2, vv2, np.ones(len(uu2)))
        assert q.info.lattice_dim == 0
    relabel = np.arange(n)
    relabel[[0, 1]] = relabel[[1, 0]]
    q = ptx.Problem.from_coo(n, relabel[uu], relabel[vv], ww)
    assert q.info.lattice_dim == 0


def test_non_lattices_are_not_detected(ptx):
    n, u, v, w = instances.erdos_renyi(400, 0.01, seed=2, pm=True)
    assert ptx.Problem.from_coo(n, u, v, w).info.lattice_dim == 0
    n, u, v, w = _lattice_edges(2, 2)  # L = 2: doubled bonds
    assert ptx.Problem.from_coo(n, u, v, w).info.lattice_dim == 0
    assert ptx.Problem.read(graph_path("G022")).info.lattice_dim == 0
