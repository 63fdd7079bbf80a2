"""Bit-exact parity of the HIP engine with the CPU oracle (colour-class order,
same counter-based random stream): spins, int energies, flips, best states,
temperature permutation, swap counters, ICM clusters and histograms.  Every
test runs through the C ABI of libptx.so on the GPU."""
import os
import sys

import numpy as np
import pytest

from conftest import REPO, graph_path

sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
import instances  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_pair(ptx, oracle, kind, **kw):
    """Returns (product Problem, oracle Graph) for a named test instance."""
    if kind.endswith(".gset") or kind.startswith("G0"):
        return ptx.Problem.read(graph_path(kind)), oracle.Graph.read(graph_path(kind))
    if kind == "pm_torus20":  # C1's instance: 2D +-J torus L=20
        n, u, v, w = instances.torus_2d(20, seed=0)
        h = None
    elif kind == "ea3d_6":
        n, u, v, w = instances.ea_3d(6, seed=kw.get("seed", 1))
        h = None
    elif kind == "field_random":
        rng = np.random.default_rng(12)
        n = 300
        m = 1500
        u = rng.integers(0, n, m).astype(np.int32)
        v = rng.integers(0, n, m).astype(np.int32)
        keep = u != v
        u, v = u[keep], v[keep]
        w = rng.integers(-3, 4, len(u)).astype(np.float64)
        h = rng.integers(-2, 3, n).astype(np.float64)
    elif kind == "er_small":
        n, u, v, w = instances.erdos_renyi(500, 0.02, seed=3, pm=True)
        h = None
    else:
        raise KeyError(kind)
    return ptx.Problem.from_coo(n, u, v, w, h), oracle.Graph.from_coo(n, u, v, w, h)


def compare(eng, oe, inst=0, what=""):
    sp = eng.spins(inst)
    osp = oe.spins()
    assert np.array_equal(sp, osp), f"{what}: spins differ in {np.sum(sp != osp)} places"
    assert np.array_equal(eng.energies(inst), oe.energies()), what
    assert np.array_equal(eng.slot2rep(inst), oe.slot2rep()), what
    assert np.array_equal(eng.flips(inst), oe.flips()), what
    be, bs, bst = eng.best(inst)
    obe, obs, obst = oe.best()
    assert np.array_equal(be, obe) and np.array_equal(bs, obs), what
    assert np.array_equal(bst, obst), what
    a, s, n = eng.swap_stats(inst)
    oa, os_, on = oe.swap_stats()
    assert np.array_equal(a, oa) and np.array_equal(s, os_) and n == on, what


# ------------------------------------------------------------- primitives ---
def test_device_philox_known_answers(ptx, oracle):
    ctr = np.array([[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]], np.uint32)
    keys = [[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]]
    for i in range(3):
        got = ptx.device_philox(ctr[i:i + 1], keys[i])[0]
        assert list(got) == list(oracle.philox(ctr[i], keys[i]))
    rng = np.random.default_rng(0)
    c = rng.integers(0, 2 ** 32, (4096, 4), dtype=np.uint64).astype(np.uint32)
    got = ptx.device_philox(c, [123, 456])
    for i in range(0, 4096, 97):
        assert list(got[i]) == list(oracle.philox(c[i], [123, 456]))


def test_device_fp64_bitwise(ptx, oracle):
    L = oracle.lib()
    rng = np.random.default_rng(1)
    x = np.concatenate([-rng.random(50000) * 60, -rng.random(5000) * 745, [0.0, -745.0, -708.3]])
    e, _ = ptx.device_fp64(x)
    bad = [i for i in range(len(x)) if e[i] != L.or_exp(float(x[i]))]
    assert not bad, (len(bad), x[bad[:3]])
    u = np.concatenate([rng.integers(1, 2 ** 32, 50000) * 2.0 ** -32, [1.0, 0.5, 2 ** -32]])
    _, lg = ptx.device_fp64(u)
    bad = [i for i in range(len(u)) if lg[i] != L.or_log(float(u[i]))]
    assert not bad, (len(bad), u[bad[:3]])


# ------------------------------------------------------------------ sweeps --
SWEEP_CASES = ["lattice_2d_5.gset", "lattice_2d_20.gset", "pm_torus20", "G022", "field_random", "ea3d_6"]


MODES = [1, 2, 3]  # split kernels, fused per-instance kernel, cooperative per-replica kernel


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kind", SWEEP_CASES)
def test_init_and_sweep_parity(ptx, oracle, kind, mode):
    p, g = make_pair(ptx, oracle, kind)
    nt = 6
    betas = np.linspace(0.3, 2.0, nt)
    say("case", kind, "n", p.n, "mode", mode)
    eng = ptx.Engine(p, betas, seed=17, flags=ptx.FLAG_NO_EXCHANGE, mode=mode)
    assert eng.info().mode == mode
    oe = oracle.Engine(g, betas, chains=2, seed=17, flags=oracle.FLAG_NO_EXCHANGE)
    compare(eng, oe, what="init")
    for s in range(4):
        eng.step(1)
        oe.step(1)
        compare(eng, oe, what=f"step {s}")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("sweeps", [2, 3])
def test_multiple_sweeps_per_step(ptx, oracle, sweeps, mode):
    p, g = make_pair(ptx, oracle, "lattice_2d_20.gset")
    betas = np.linspace(0.4, 1.2, 5)
    eng = ptx.Engine(p, betas, seed=2, sweeps=sweeps, flags=ptx.FLAG_NO_EXCHANGE, mode=mode)
    oe = oracle.Engine(g, betas, chains=2, seed=2, sweeps=sweeps, flags=oracle.FLAG_NO_EXCHANGE)
    eng.step(3)
    oe.step(3)
    compare(eng, oe, what="sweeps")


# ------------------------------------------------------- full PT steps ------
FULL_CASES = [("lattice_2d_20.gset", 8, 0.5, 3.0, 20), ("pm_torus20", 32, 0.5, 3.0, 12),
              ("G022", 16, 0.6, 2.8, 6), ("field_random", 10, 0.5, 3.0, 10), ("er_small", 12, 0.4, 2.5, 10)]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kind,nt,T0,T1,steps", FULL_CASES)
@pytest.mark.parametrize("icm", [False, True])
def test_full_step_parity(ptx, oracle, kind, nt, T0, T1, steps, icm, mode):
    p, g = make_pair(ptx, oracle, kind)
    betas = oracle.beta_linspace(T0, T1, nt, 0)
    say("case", kind, "nt", nt, "icm", icm, "mode", mode)
    flags = ptx.FLAG_ICM if icm else 0
    nc = -1 if kind != "field_random" else nt // 2
    eng = ptx.Engine(p, betas, seed=5, flags=flags, nc=nc, mode=mode, steps_per_launch=3)
    oe = oracle.Engine(g, betas, chains=2, seed=5, flags=(oracle.FLAG_ICM if icm else 0), nc=nc)
    if icm:
        assert eng.instance_info().nc == oe.nc
    for s in range(steps):
        eng.step(1)
        oe.step(1)
        compare(eng, oe, what=f"{kind} step {s}")
    # bulk stepping (no host reads in between) ends in the same state
    eng.step(25)
    oe.step(25)
    compare(eng, oe, what=f"{kind} bulk")


@pytest.mark.parametrize("mode", MODES)
def test_icm_all_slots_parity(ptx, oracle, mode):
    """ICM on every slot (nc = nt), including global flips of replica 1."""
    p, g = make_pair(ptx, oracle, "G039")
    nt = 8
    betas = oracle.beta_linspace(0.5, 2.0, nt, 0)
    eng = ptx.Engine(p, betas, seed=33, flags=ptx.FLAG_ICM, nc=nt, mode=mode)
    oe = oracle.Engine(g, betas, chains=2, seed=33, flags=oracle.FLAG_ICM, nc=nt)
    for s in range(12):
        eng.step(1)
        oe.step(1)
        compare(eng, oe, what=f"step {s}")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("dim,L", [(2, 33), (2, 40), (2, 64), (2, 80), (2, 100), (2, 129), (3, 5), (3, 12),
                                   (3, 24)])
def test_lattice_cluster_growth_parity(ptx, oracle, dim, L, mode):
    """Periodic lattices grow clusters as bit sets (split and cooperative
    kernels): 2D rows of 2, 3 or 4 words held whole in registers, one or two
    rows per lane (L = 33 .. 100: carry-add row fills, rows y +- 1 by DPP
    shifts), rows of one word per lane (3D L = 5), words in registers and
    LDS (3D L = 12), or updated in place in LDS (2D L = 129, 3D L = 24); odd
    L with three colour classes.  The oracle's breadth-first clusters bit
    for bit, with ICM on every slot (global flips included)."""
    if dim == 2:
        n, u, v, w = instances.torus_2d(L, seed=L)
    else:
        n, u, v, w = instances.ea_3d(L, seed=L)
    p, g = ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)
    assert (p.info.lattice_dim, p.info.lattice_L) == (dim, L)
    nt = 8
    betas = oracle.beta_linspace(0.4, 3.0, nt, 0)
    say("lattice", dim, L, "mode", mode)
    eng = ptx.Engine(p, betas, seed=L, flags=ptx.FLAG_ICM, nc=nt, mode=mode, steps_per_launch=2)
    assert eng.info().mode == mode
    oe = oracle.Engine(g, betas, chains=2, seed=L, flags=oracle.FLAG_ICM, nc=nt)
    for s_ in range(6):
        eng.step(1)
        oe.step(1)
        compare(eng, oe, what=f"L={L} dim={dim} step {s_}")
    eng.step(10)
    oe.step(10)
    compare(eng, oe, what=f"L={L} dim={dim} bulk")


@pytest.mark.parametrize("dim,L,pm", [(2, 4, True), (2, 6, True), (2, 10, False), (2, 32, True), (2, 64, True),
                                      (2, 66, True), (2, 100, False), (2, 128, True), (3, 4, True), (3, 6, True),
                                      (3, 10, False), (3, 16, True)])
@pytest.mark.parametrize("icm", [False, True])
def test_multispin_lattice_sweep_parity(ptx, oracle, dim, L, pm, icm):
    """Split-mode sweeps of even periodic lattices with couplings +-1 run the
    multispin sweep on the state bits (rows of one or two words: L <= 64 and
    L = 66 .. 128; L = 4; class boundaries inside a Philox block: 2D L = 6,
    10), bit for bit with the oracle's site-by-site sweep; ferromagnets (pm
    False, w = -1) and +-J couplings; two sweeps per step for the last part."""
    if dim == 2:
        n, u, v, w = instances.torus_2d(L, seed=L, pm=pm)
    else:
        n, u, v, w = instances.ea_3d(L, seed=L)
        if not pm:
            w = -np.ones_like(w)
    p, g = ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)
    nt = 8
    betas = oracle.beta_linspace(0.3, 3.0, nt, 0)
    say("multispin lattice", dim, L, "pm", pm, "icm", icm)
    flags = ptx.FLAG_ICM if icm else 0
    eng = ptx.Engine(p, betas, seed=3 * L, flags=flags, mode=1, waves_per_group=1)
    assert eng.instance_info().lattice_sweep == dim
    oe = oracle.Engine(g, betas, chains=2, seed=3 * L, flags=(oracle.FLAG_ICM if icm else 0))
    for s_ in range(4):
        eng.step(1)
        oe.step(1)
        compare(eng, oe, what=f"L={L} dim={dim} step {s_}")
    eng.step(20)
    oe.step(20)
    compare(eng, oe, what=f"L={L} dim={dim} bulk")
    eng2 = ptx.Engine(p, betas, seed=L, flags=flags, mode=1, waves_per_group=2, sweeps=2)
    oe2 = oracle.Engine(g, betas, chains=2, seed=L, sweeps=2, flags=(oracle.FLAG_ICM if icm else 0))
    eng2.step(6)
    oe2.step(6)
    compare(eng2, oe2, what=f"L={L} dim={dim} two sweeps")


@pytest.mark.parametrize("wpg", [2, 4, 8])
def test_multispin_workgroup_sweep_parity(ptx, oracle, wpg):
    """Multispin sweeps with several waves (replicas) per workgroup:
    bit-exact with the oracle for a batch of 3D EA L = 4, 6, 8 and 2D
    L = 6, 10, 32 lattices, with cluster moves, on ladders of 5, 7 and 3
    slots (10, 14 and 6 replicas: workgroups whose last waves have no
    replica), and with two sweeps per step."""
    specs = [(3, 4), (3, 6), (3, 8), (2, 6), (2, 10), (2, 32)]
    pairs = []
    for dim, L in specs:
        n, u, v, w = instances.ea_3d(L, seed=L) if dim == 3 else instances.torus_2d(L, seed=L, pm=True)
        pairs.append((ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)))
    nts = [5, 7, 3, 5, 7, 3]
    ladders = [oracle.beta_linspace(0.3, 3.0, nt, 0) for nt in nts]
    for sweeps in (1, 2):
        say("multispin workgroup sweep W", wpg, "sweeps", sweeps)
        eng = ptx.Engine([pp for pp, _ in pairs], ladders, seed=11 + sweeps, flags=ptx.FLAG_ICM, mode=1,
                         waves_per_group=wpg, sweeps=sweeps)
        assert eng.info().waves_per_group == wpg
        assert all(eng.instance_info(i).lattice_sweep == specs[i][0] for i in range(len(specs)))
        oes = [oracle.Engine(g, lad, chains=2, seed=11 + sweeps, instance=i, sweeps=sweeps, flags=oracle.FLAG_ICM)
               for i, ((_, g), lad) in enumerate(zip(pairs, ladders))]
        for s_ in range(3):
            eng.step(1)
            for oe in oes:
                oe.step(1)
            for i, oe in enumerate(oes):
                compare(eng, oe, inst=i, what=f"W={wpg} sweeps={sweeps} inst {i} step {s_}")
        eng.step(15)
        for oe in oes:
            oe.step(15)
        for i, oe in enumerate(oes):
            compare(eng, oe, inst=i, what=f"W={wpg} sweeps={sweeps} inst {i} bulk")
        eng.close()


def test_multispin_lattice_sweep_only_where_it_applies(ptx, oracle):
    """Odd L (three colour classes), a field, or couplings other than +-1 keep
    the gather sweep; the engine reports which instances take which."""
    cases = []
    n, u, v, w = instances.torus_2d(9, seed=1)
    cases.append(("odd L", (n, u, v, w, None), 0))
    n, u, v, w = instances.torus_2d(8, seed=1)
    cases.append(("field", (n, u, v, w, np.where(np.arange(n) % 7 == 0, 1.0, 0.0)), 0))
    cases.append(("weights 2", (n, u, v, 2 * w, None), 0))
    cases.append(("even L", (n, u, v, w, None), 2))
    probs = [ptx.Problem.from_coo(*c[1]) for c in cases]
    graphs = [oracle.Graph.from_coo(*c[1]) for c in cases]
    betas = oracle.beta_linspace(0.4, 2.5, 6, 0)
    eng = ptx.Engine(probs, [betas] * len(probs), seed=9, mode=1)
    for i, (what, _, want) in enumerate(cases):
        assert eng.instance_info(i).lattice_sweep == want, what
    eng.step(8)
    for i, g in enumerate(graphs):
        oe = oracle.Engine(g, betas, chains=2, seed=9, instance=i)
        oe.step(8)
        compare(eng, oe, inst=i, what=cases[i][0])


@pytest.mark.parametrize("kind,nc", [("G039", 8), ("pm_torus20", -1), ("field_random", 4)])
@pytest.mark.parametrize("icm_waves", [1, 2, 4])
def test_icm_waves_independence(ptx, oracle, kind, nc, icm_waves):
    """Split mode: a cluster grown by 1, 2 or 4 waves (queue order differs,
    the component does not) gives the oracle's states and energies."""
    p, g = make_pair(ptx, oracle, kind)
    betas = oracle.beta_linspace(0.5, 2.0, 8, 0)
    eng = ptx.Engine(p, betas, seed=41, flags=ptx.FLAG_ICM, nc=nc, mode=ptx.MODE_SPLIT, icm_waves=icm_waves)
    assert eng.info().icm_waves == icm_waves
    oe = oracle.Engine(g, betas, chains=2, seed=41, flags=oracle.FLAG_ICM, nc=nc)
    for s in range(10):
        eng.step(1)
        oe.step(1)
        compare(eng, oe, what=f"{kind} icm_waves={icm_waves} step {s}")


@pytest.mark.parametrize("mode,wpg,spl,lanes", [(1, 1, 0, 0), (1, 2, 0, 0), (1, 4, 0, 0), (1, 8, 0, 0), (2, 4, 1, 0),
                                                (2, 8, 7, 0), (2, 16, 16, 0), (3, 4, 5, 1), (3, 8, 16, 1),
                                                (3, 4, 1, 4), (3, 8, 4, 4)])
def test_launch_shape_independence(ptx, oracle, mode, wpg, spl, lanes):
    p, g = make_pair(ptx, oracle, "pm_torus20")
    betas = oracle.beta_linspace(0.5, 3.0, 16, 0)
    eng = ptx.Engine(p, betas, seed=9, flags=ptx.FLAG_ICM, waves_per_group=wpg, mode=mode, steps_per_launch=spl,
                     spins_per_lane=lanes)
    assert eng.info().mode == mode
    oe = oracle.Engine(g, betas, chains=2, seed=9, flags=oracle.FLAG_ICM)
    eng.step(15)
    oe.step(15)
    compare(eng, oe, what=f"W={wpg}")


@pytest.mark.parametrize("batch", ["unit", "wide"])
@pytest.mark.parametrize("wpg", [2, 4, 8])
def test_multiwave_gather_sweep_parity(ptx, oracle, batch, wpg):
    """Split gather sweeps with several waves (replicas of one instance) per
    workgroup: bit-exact with the oracle for rows of one chunk (an odd
    torus), of several chunks (G022), with the sign table (+-J ER), in the
    32-bit record format with a field, and with workgroups whose last waves
    have no replica (2 nt not a multiple of the wave count)."""
    if batch == "unit":
        pairs = [make_pair(ptx, oracle, k) for k in ["G022", "er_small"]]
        n, u, v, w = instances.torus_2d(9, seed=4)
        pairs.append((ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)))
    else:
        pairs = [make_pair(ptx, oracle, k) for k in ["field_random", "G022"]]
    nts = [5, 7, 3][: len(pairs)]
    ladders = [oracle.beta_linspace(0.5, 3.0, nt, 0) for nt in nts]
    say("multiwave gather", batch, "W", wpg)
    eng = ptx.Engine([pp for pp, _ in pairs], ladders, seed=21, flags=ptx.FLAG_ICM, mode=1, waves_per_group=wpg,
                     sliced_sweep=-1)
    assert eng.info().waves_per_group == wpg
    oes = [oracle.Engine(g, lad, chains=2, seed=21, instance=i, flags=oracle.FLAG_ICM)
           for i, ((_, g), lad) in enumerate(zip(pairs, ladders))]
    for s_ in range(3):
        eng.step(1)
        for oe in oes:
            oe.step(1)
        for i, oe in enumerate(oes):
            compare(eng, oe, inst=i, what=f"{batch} W={wpg} inst {i} step {s_}")
    eng.step(12)
    for oe in oes:
        oe.step(12)
    for i, oe in enumerate(oes):
        compare(eng, oe, inst=i, what=f"{batch} W={wpg} inst {i} bulk")


@pytest.mark.parametrize("sliced,sweeps", [(0, 1), (0, 2), (1, 1), (2, 2), (4, 2), (8, 1), (-1, 1)])
def test_replica_sliced_sweep_parity(ptx, oracle, sliced, sweeps):
    """Split mode with the replica-sliced sweep (32 replicas per workgroup of
    1, 2, 4 or 8 waves, neighbour counts bit-sliced) and without it:
    bit-exact with the oracle for the instances it takes (G022 with replica
    blocks of 32 + 32 + 16, a +-J ER graph with 10 replicas, an odd torus
    whose rows are one chunk, a +-J G(5000, 0.004) with 32 replicas) next to
    ones it leaves to the other sweeps (a field, degree above 63, an even
    torus on the multispin sweep), all in the 16-bit record format.  Chosen
    automatically (sliced = 0) it takes only instances with 32 replicas or
    more whose colour classes average 512 positions or more: here the
    G(5000, 0.004) graph."""
    pairs = [make_pair(ptx, oracle, k) for k in ["G022", "er_small"]]
    n, u, v, w = instances.torus_2d(9, seed=4)
    pairs.append((ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)))
    n, u, v, w = instances.erdos_renyi(200, 0.05, seed=6, pm=True)
    h = np.where(np.arange(n) % 5 == 0, 1.0, 0.0)
    pairs.append((ptx.Problem.from_coo(n, u, v, w, h), oracle.Graph.from_coo(n, u, v, w, h)))
    n, u, v, w = instances.erdos_renyi(300, 0.3, seed=5, pm=True)
    pairs.append((ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)))
    n, u, v, w = instances.torus_2d(8, seed=2)
    pairs.append((ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)))
    n, u, v, w = instances.erdos_renyi(5000, 0.004, seed=7, pm=True)
    pairs.append((ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)))
    want = {0: [0, 0, 0, 0, 0, 0, 1], -1: [0] * 7}.get(sliced, [1, 1, 1, 0, 0, 0, 1])
    nts = [40, 5, 7, 6, 6, 6, 16]
    ladders = [oracle.beta_linspace(0.4, 3.0, nt, 0) for nt in nts]
    say("replica-sliced", sliced, "sweeps", sweeps)
    eng = ptx.Engine([pp for pp, _ in pairs], ladders, seed=23, sweeps=sweeps, flags=ptx.FLAG_ICM, mode=1,
                     sliced_sweep=sliced)
    assert [eng.instance_info(i).sliced_sweep for i in range(len(pairs))] == want
    oes = [oracle.Engine(g, lad, chains=2, seed=23, sweeps=sweeps, instance=i, flags=oracle.FLAG_ICM)
           for i, ((_, g), lad) in enumerate(zip(pairs, ladders))]
    for s_ in range(3):
        eng.step(1)
        for oe in oes:
            oe.step(1)
        for i, oe in enumerate(oes):
            compare(eng, oe, inst=i, what=f"sliced={sliced} inst {i} step {s_}")
    eng.step(20)
    for oe in oes:
        oe.step(20)
    for i, oe in enumerate(oes):
        compare(eng, oe, inst=i, what=f"sliced={sliced} inst {i} bulk")


@pytest.mark.parametrize("kind,steps,per_launch", [("G022", 300, 500), ("lattice_2d_20.gset", 300, 500),
                                                    ("er_small", 300, 500), ("G022", 60, 7)])
def test_cooperative_best_state_over_long_launches(ptx, oracle, kind, steps, per_launch):
    """The cooperative kernel without cluster moves keeps a replica's best
    state as a log of the flips after the last new best, built into the best
    bytes at the end of a launch or before a class that could overflow the
    log (long launches with rare new bests); best energies, sweeps and states
    stay bit-exact with the oracle's after every part of the run."""
    p, g = make_pair(ptx, oracle, kind)
    nt = 11
    betas = oracle.beta_linspace(0.4, 3.0, nt, 0)
    say("coop best", kind, steps, per_launch)
    eng = ptx.Engine(p, betas, seed=41, flags=0, mode=3, steps_per_launch=per_launch)  # no cluster moves
    oe = oracle.Engine(g, betas, chains=2, seed=41, flags=0)
    assert eng.instance_info().nc == 0  # the kernel without cluster-move code
    for part in (steps // 3, steps - steps // 3):
        eng.step(part)
        oe.step(part)
        compare(eng, oe, what=f"{kind} after {part} more steps")


@pytest.mark.parametrize("kind", ["G022", "field_random", "ea3d_6"])
def test_cooperative_16_waves_parity(ptx, oracle, kind):
    """Cooperative kernel with 16 waves per replica (no ICM slot; 4 lanes
    per position on classes of at most 256 positions, 2 on classes of at
    most 512) against the oracle."""
    p, g = make_pair(ptx, oracle, kind)
    betas = oracle.beta_linspace(0.5, 3.0, 10, 0)
    say("case", kind, "16 waves")
    eng = ptx.Engine(p, betas, seed=23, flags=0, mode=ptx.MODE_COOP, waves_per_group=16, steps_per_launch=4)
    assert eng.info().waves_per_group == 16
    oe = oracle.Engine(g, betas, chains=2, seed=23, flags=0)
    for s_ in range(6):
        eng.step(1)
        oe.step(1)
        compare(eng, oe, what=f"{kind} step {s_}")
    eng.step(20)
    oe.step(20)
    compare(eng, oe, what=f"{kind} bulk")


def test_cooperative_16_waves_fall_back_to_8_with_cluster_moves(ptx, oracle):
    p, g = make_pair(ptx, oracle, "pm_torus20")
    betas = oracle.beta_linspace(0.5, 3.0, 12, 0)
    eng = ptx.Engine(p, betas, seed=4, flags=ptx.FLAG_ICM, mode=ptx.MODE_COOP, waves_per_group=16)
    assert eng.info().waves_per_group == 8
    oe = oracle.Engine(g, betas, chains=2, seed=4, flags=oracle.FLAG_ICM)
    eng.step(10)
    oe.step(10)
    compare(eng, oe, what="16 -> 8 waves")


def test_write_spins_recomputes_energy(ptx, oracle):
    p, g = make_pair(ptx, oracle, "lattice_2d_20.gset")
    betas = np.linspace(0.5, 1.5, 4)
    eng = ptx.Engine(p, betas, seed=1)
    rng = np.random.default_rng(3)
    s = rng.integers(0, 2, (8, p.n)).astype(np.uint8)
    eng.set_spins(s)
    assert np.array_equal(eng.spins(), s)
    assert np.array_equal(eng.energies(), [g.energy(x) for x in s])


# ---------------------------------------------------------- pilot / hist ----
@pytest.mark.parametrize("mode", MODES)
def test_pilot_histograms_match_oracle_logs(ptx, oracle, mode):
    p, g = make_pair(ptx, oracle, "lattice_2d_20.gset")
    nt, burn, ns = 10, 100, 400
    hb, bs, be = ptx.pilot_betas(0.6, 2.8, nt)
    eng = ptx.Engine(p, hb, seed=21, chains=1, flags=ptx.FLAG_PILOT | ptx.FLAG_HISTOGRAM, hist_start=burn, mode=mode)
    eng.step(burn + ns)
    counts, e_base = eng.histograms()
    logs = oracle.gather_samples(g, hb, 21, burn, ns)
    assert counts.sum() == ns * nt
    for j in range(nt):
        want = np.bincount((logs[:, j] - e_base) // 2, minlength=counts.shape[1])
        assert np.array_equal(counts[j], want), j
    lad = ptx.ladder_from_histograms(counts, e_base, hb, bs, be, mode=ptx.LADDER_EXACT)
    olad, _, _ = oracle.get_beta_from_logs(logs, hb, 2.0, 0.1, 0.4, bs, be)
    assert np.array_equal(lad, olad)


# ---------------------------------------------------------------- batched ---
@pytest.mark.parametrize("mode,csr", [(1, 0), (2, 1), (2, 2), (3, 0)])
@pytest.mark.parametrize("with_field", [False, True])
def test_batched_instances_match_single_oracles(ptx, oracle, mode, csr, with_field):
    """csr 1: adjacency staged in LDS; csr 2: the fused kernel reads it through
    L2; mode 3: one cooperative launch for the whole batch."""
    kinds = ["lattice_2d_10.gset", "ea3d_6", "pm_torus20", "lattice_2d_5.gset", "er_small"]
    if with_field:
        kinds.append("field_random")  # forces the 32-bit adjacency format in the fused kernel
    pairs = [make_pair(ptx, oracle, k) for k in kinds]
    nts = [4, 8, 6, 3, 5, 7][: len(kinds)]
    ladders = [oracle.beta_linspace(0.5, 3.0, nt, 0) for nt in nts]
    eng = ptx.Engine([pp for pp, _ in pairs], ladders, seed=77, flags=ptx.FLAG_ICM, mode=mode, steps_per_launch=2,
                     csr_placement=csr)
    if mode == 2:
        assert eng.info().csr_in_lds == (1 if csr == 1 else 0)
    oes = [oracle.Engine(g, lad, chains=2, seed=77, instance=i, flags=oracle.FLAG_ICM)
           for i, ((_, g), lad) in enumerate(zip(pairs, ladders))]
    for s in range(6):
        eng.step(1)
        for oe in oes:
            oe.step(1)
        for i, oe in enumerate(oes):
            compare(eng, oe, inst=i, what=f"inst {i} step {s}")


def test_cooperative_batch_reads_adjacency_through_l2_when_lds_is_full(ptx, oracle):
    """Six G022-sized ER graphs: one instance's adjacency fills most of a
    CU's LDS, so 288 resident workgroups need the form that reads the
    adjacency through L2 -- bit-exact with the oracle like the LDS form."""
    pairs = []
    for k in range(6):
        n, u, v, w = instances.erdos_renyi(2000, 0.01, seed=40 + k, pm=k % 2 == 1)
        pairs.append((ptx.Problem.from_coo(n, u, v, w), oracle.Graph.from_coo(n, u, v, w)))
    lad = oracle.beta_linspace(0.5, 3.0, 24, 0)
    eng = ptx.Engine([p for p, _ in pairs], [lad] * 6, seed=13, flags=ptx.FLAG_ICM, nc=6, mode=ptx.MODE_COOP,
                     steps_per_launch=2)
    oes = [oracle.Engine(g, lad, chains=2, seed=13, instance=i, flags=oracle.FLAG_ICM, nc=6)
           for i, (_, g) in enumerate(pairs)]
    for s_ in range(3):
        eng.step(1)
        for oe in oes:
            oe.step(1)
        for i, oe in enumerate(oes):
            compare(eng, oe, inst=i, what=f"inst {i} step {s_}")
    assert eng.info().csr_in_lds == 0


def test_batched_3d_ea_invariants(ptx, oracle):
    """A C3-shaped batch (uneven 3D EA instances, 32 temperatures, ICM):
    tracked energies equal recomputed ones, permutations stay within chains."""
    rng = np.random.default_rng(0)
    Ls = rng.choice([6, 8, 10, 12], size=48)
    probs = []
    graphs = []
    for i, L in enumerate(Ls):
        n, u, v, w = instances.ea_3d(int(L), seed=1000 + i)
        probs.append(ptx.Problem.from_coo(n, u, v, w))
        graphs.append(oracle.Graph.from_coo(n, u, v, w))
    lad = oracle.beta_linspace(0.2, 3.0, 32, 1)
    eng = ptx.Engine(probs, [lad] * len(probs), seed=3, flags=ptx.FLAG_ICM, mode=ptx.MODE_FUSED)
    assert eng.info().mode == 2
    eng.step(40)
    for i in range(0, len(probs), 7):
        sp = eng.spins(i)
        assert np.array_equal(eng.energies(i), [graphs[i].energy(x) for x in sp])
        s2r = eng.slot2rep(i)
        assert set(s2r[:32].tolist()) == set(range(32)) and set(s2r[32:].tolist()) == set(range(32, 64))
        a, s, nsw = eng.swap_stats(i)
        assert a.sum() == 40 * 2 * 31 and s.sum() == nsw
    # one instance checked bit for bit against the oracle
    oe = oracle.Engine(graphs[5], lad, chains=2, seed=3, instance=5, flags=oracle.FLAG_ICM)
    oe.step(40)
    compare(eng, oe, inst=5, what="C3 instance 5")
