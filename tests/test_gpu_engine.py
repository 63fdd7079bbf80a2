"""Engine features that must not change results: hipGraph replay of the
fused kernel and the launch-shape autotuner (the random streams are indexed
by position, replica and step, so any launch shape gives the same bits)."""
import os
import sys

import numpy as np
import pytest

from conftest import REPO

sys.path.insert(0, os.path.join(REPO, "pt-icm_amd"))
import instances  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _batch(ptx, k=12, seed=0):
    _, coos = instances.ea_batch(k, seed, sizes=(4, 6))
    probs = [ptx.Problem.from_coo(*c) for c in coos]
    lad = ptx.beta_linspace(0.2, 3.0, 12, True)
    return probs, [lad] * len(probs)


def _state(eng, k):
    return [(eng.spins(i), eng.energies(i), eng.slot2rep(i), eng.flips(i), eng.best(i)[:2]) for i in range(k)]


def _same(a, b):
    for x, y in zip(a, b):
        for u, v in zip(x[:4], y[:4]):
            assert np.array_equal(u, v)
        assert np.array_equal(x[4][0], y[4][0]) and np.array_equal(x[4][1], y[4][1])


def test_graph_replay_matches_eager_launches(ptx):
    probs, lads = _batch(ptx)
    eager = ptx.Engine(probs, lads, seed=5, flags=ptx.FLAG_ICM, mode=ptx.MODE_FUSED, steps_per_launch=2)
    graph = ptx.Engine(probs, lads, seed=5, flags=ptx.FLAG_ICM, mode=ptx.MODE_FUSED, steps_per_launch=2,
                       use_graph=3)
    graph.timing(True)
    done = 0
    for chunk in (1, 13, 6, 7, 12):  # graph replays of 6 steps, eager remainders, host reads between them
        eager.step(chunk)
        graph.step(chunk)
        done += chunk
        say("steps", done)
        _same(_state(eager, len(probs)), _state(graph, len(probs)))
    t = graph.timing()
    # launches: 1 eager | 2 replays of 3 + 1 eager | 1 replay | 1 replay + 1 eager | 2 replays
    assert t.fused_launches == 1 + 7 + 3 + 4 + 6


def test_graph_replay_of_split_steps_matches_eager(ptx):
    probs, lads = _batch(ptx, k=10, seed=7)
    eager = ptx.Engine(probs, lads, seed=8, flags=ptx.FLAG_ICM, mode=ptx.MODE_SPLIT)
    graph = ptx.Engine(probs, lads, seed=8, flags=ptx.FLAG_ICM, mode=ptx.MODE_SPLIT, use_graph=4)
    assert eager.info().mode == 1 and graph.info().mode == 1
    for chunk in (3, 9, 4, 1, 8):  # replays of 4 steps and eager remainders, reads in between
        eager.step(chunk)
        graph.step(chunk)
        _same(_state(eager, len(probs)), _state(graph, len(probs)))
    graph.set_active(np.arange(len(probs)) % 2 == 0)  # new work lists: the graph is rebuilt
    eager.set_active(np.arange(len(probs)) % 2 == 0)
    eager.step(8)
    graph.step(8)
    _same(_state(eager, len(probs)), _state(graph, len(probs)))


def test_auto_mode_runs_batches_on_the_split_kernels(ptx):
    probs, lads = _batch(ptx, k=6, seed=9)
    assert ptx.Engine(probs, lads, seed=1, flags=ptx.FLAG_ICM).info().mode == 1


def test_autotuned_shape_gives_identical_results(ptx):
    probs, lads = _batch(ptx, k=8, seed=1)
    kw, trials = ptx.autotune(probs, lads, seed=3, steps=8, flags=ptx.FLAG_ICM)
    say("autotune", kw, [(t["mode"], t["waves_per_group"], t["csr_placement"], round(t["us_per_step"], 1))
                         for t in trials])
    assert len(trials) >= 3 and kw["mode"] in (1, 2)
    assert min(t["us_per_step"] for t in trials) > 0
    ref = ptx.Engine(probs, lads, seed=3, flags=ptx.FLAG_ICM, mode=ptx.MODE_SPLIT, waves_per_group=1)
    tuned = ptx.Engine(probs, lads, seed=3, flags=ptx.FLAG_ICM, **kw)
    ref.step(10)
    tuned.step(10)
    _same(_state(ref, len(probs)), _state(tuned, len(probs)))


def test_autotune_single_instance_includes_cooperative(ptx):
    p = ptx.Problem.from_coo(*instances.torus_2d(16, seed=2))
    lad = ptx.beta_linspace(0.5, 3.0, 10, False)
    kw, trials = ptx.autotune(p, lad, seed=1, steps=16, flags=ptx.FLAG_ICM)
    assert any(t["mode"] == 3 for t in trials)
    a = ptx.Engine(p, lad, seed=1, flags=ptx.FLAG_ICM, **kw)
    b = ptx.Engine(p, lad, seed=1, flags=ptx.FLAG_ICM, mode=ptx.MODE_SPLIT)
    a.step(9)
    b.step(9)
    _same(_state(a, 1), _state(b, 1))


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_retired_instances_freeze_and_the_rest_match(ptx, mode):
    probs, lads = _batch(ptx, k=12, seed=2)
    ref = ptx.Engine(probs, lads, seed=4, flags=ptx.FLAG_ICM, mode=mode, steps_per_launch=1)
    eng = ptx.Engine(probs, lads, seed=4, flags=ptx.FLAG_ICM, mode=mode, steps_per_launch=1)
    ref.step(5)
    eng.step(5)
    frozen = _state(eng, len(probs))
    mask = np.array([i % 3 != 1 for i in range(len(probs))])
    eng.set_active(mask)
    ref.step(7)
    eng.step(7)
    now, want = _state(eng, len(probs)), _state(ref, len(probs))
    for i in range(len(probs)):
        if mask[i]:
            _same([want[i]], [now[i]])   # the active ones run exactly as in the full batch
        else:
            _same([frozen[i]], [now[i]])  # the retired ones did not move
    eng.set_active(np.ones(len(probs), bool))
    eng.step(3)
    for i, p in enumerate(probs):  # reactivated instances keep consistent energies
        sp = eng.spins(i)
        assert np.array_equal(eng.energies(i), [p.energy(x) for x in sp])
    with pytest.raises(ptx.PtxError):
        eng.set_active(np.zeros(len(probs), bool))
        eng.set_spins(eng.spins(0), 0)


def test_small_batch_coop_switches_without_changing_results(ptx):
    """A split-mode batch too large for one cooperative launch runs the split
    kernels; once retirement leaves few enough replicas it runs the
    cooperative kernel -- with the same trajectories as a split-only engine
    (every kernel draws from the instance's own step counter)."""
    _, coos = instances.ea_batch(96, 7, sizes=(4, 6))
    probs = [ptx.Problem.from_coo(*c) for c in coos]
    lad = ptx.beta_linspace(0.2, 3.0, 16, True)
    lads = [lad] * len(probs)
    ref = ptx.Engine(probs, lads, seed=8, flags=ptx.FLAG_ICM, mode=ptx.MODE_SPLIT)
    eng = ptx.Engine(probs, lads, seed=8, flags=ptx.FLAG_ICM, mode=ptx.MODE_SPLIT, small_batch_coop=1, use_graph=2)
    ref.step(6)
    eng.step(6)
    assert eng.info().mode == ptx.MODE_SPLIT  # 3072 workgroups: more than any MI355X keeps resident
    say("96 x 32 replicas: split")
    _same(_state(ref, len(probs)), _state(eng, len(probs)))
    mask = np.array([i % 12 == 0 for i in range(len(probs))])  # 8 x 32 replicas left
    ref.set_active(mask)
    eng.set_active(mask)
    ref.step(9)
    eng.step(9)
    assert eng.info().mode == ptx.MODE_COOP
    _same(_state(ref, len(probs)), _state(eng, len(probs)))
    for i in np.nonzero(mask)[0]:
        sp = eng.spins(i)
        assert np.array_equal(eng.energies(i), [probs[i].energy(x) for x in sp])


def test_time_to_target_warm_start_maps_slots_by_nearest_beta(ptx):
    """The time-to-target harness hands a fallback ladder the replicas'
    current states: slot t of each chain of the new engine starts from the
    state at the slot of the same chain with the nearest beta in the old one
    (after the old engine has exchanged, so its permutation is not the
    identity); energies are recomputed from the written states."""
    import ttt
    probs, _ = _batch(ptx, k=3, seed=5)
    lad_a = np.sort(ptx.beta_linspace(0.3, 2.5, 10, True))
    lad_b = np.sort(ptx.beta_linspace(0.35, 2.4, 7, True))
    src = ptx.Engine(probs, [lad_a] * 3, seed=3, flags=ptx.FLAG_ICM, mode=1)
    src.step(40)
    dst = ptx.Engine(probs[::-1], [lad_b] * 3, seed=4, flags=ptx.FLAG_ICM, mode=1)  # entry 0 is problem 2
    dst.step(3)
    ttt.warm_start(src, 2, lad_a, dst, 0, lad_b)
    sp_a, s2r_a, e_a = src.spins(2), src.slot2rep(2), src.energies(2)
    sp_b, s2r_b, e_b = dst.spins(0), dst.slot2rep(0), dst.energies(0)
    assert not np.array_equal(s2r_a, np.arange(len(s2r_a)))
    for c in range(2):
        for t in range(len(lad_b)):
            s = int(np.argmin(np.abs(lad_a - lad_b[t])))
            ra, rb = s2r_a[c * len(lad_a) + s], s2r_b[c * len(lad_b) + t]
            assert np.array_equal(sp_b[rb], sp_a[ra]), (c, t)
            assert e_b[rb] == e_a[ra], (c, t)
    src.close()
    dst.close()
    for p in probs:
        p.close()


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_copies_of_one_problem_share_tables_without_changing_results(ptx, mode):
    """Entries of a batch that hold the same Problem share its read-only
    tables (rows, adjacency, classes, thresholds for equal ladders, lattice
    tables): every entry equals the same batch built from separate Problem
    objects, bit for bit (a 2D +-J torus with lattice cluster moves and the
    multispin sweep, an odd torus, a +-J ER graph; one copy on another
    ladder)."""
    coos = [instances.torus_2d(12, seed=1), instances.torus_2d(9, seed=2),
            instances.erdos_renyi(300, 0.03, seed=3, pm=True)]
    shared = [ptx.Problem.from_coo(*c) for c in coos]
    lad = np.sort(ptx.beta_linspace(0.4, 2.8, 8, True))
    lad2 = np.sort(ptx.beta_linspace(0.5, 2.5, 8, True))
    order = [0, 1, 2, 0, 2, 1, 0]
    lads = [lad] * 6 + [lad2]
    separate = [ptx.Problem.from_coo(*coos[k]) for k in order]
    say("copies mode", mode)
    a = ptx.Engine([shared[k] for k in order], lads, seed=9, flags=ptx.FLAG_ICM, mode=mode)
    b = ptx.Engine(separate, lads, seed=9, flags=ptx.FLAG_ICM, mode=mode)
    a.step(30)
    b.step(30)
    _same(_state(a, len(order)), _state(b, len(order)))
    # copies run their own random streams
    assert not np.array_equal(a.spins(0), a.spins(3))
    a.close()
    b.close()


@pytest.mark.parametrize("waves", [0, 4])
def test_replica_sliced_sweep_with_pilot_histograms_observables_and_checkpoints(ptx, waves):
    """The replica-sliced sweep equals the per-replica sweep bit for bit in a
    pilot run with histograms (one chain, every pair exchanged), with
    observables, and across a checkpoint taken by one engine and resumed by
    the other (+-J ER graphs of 5,000 and 600 vertices; waves = 0 takes the
    engine's own choice, the larger graph only, 4 forces the sliced sweep on
    both)."""
    coos = [instances.erdos_renyi(5000, 0.004, seed=2, pm=True), instances.erdos_renyi(600, 0.03, seed=4, pm=True)]
    probs = [ptx.Problem.from_coo(*c) for c in coos]
    pil = [ptx.pilot_betas(0.5, 3.0, 40, False)[0]] * 2
    kw = dict(seed=13, chains=1, flags=ptx.FLAG_PILOT | ptx.FLAG_HISTOGRAM, mode=ptx.MODE_SPLIT)
    a = ptx.Engine(probs, pil, sliced_sweep=waves, **kw)
    b = ptx.Engine(probs, pil, sliced_sweep=-1, **kw)
    assert a.instance_info(0).sliced_sweep == 1 and b.instance_info(0).sliced_sweep == 0
    a.step(40)
    b.step(40)
    for i in range(2):
        assert np.array_equal(a.histograms(i)[0], b.histograms(i)[0])
    _same(_state(a, 2), _state(b, 2))
    a.close()
    b.close()
    lads = [np.sort(ptx.beta_linspace(0.5, 3.0, 24, False))] * 2
    kw = dict(seed=17, flags=ptx.FLAG_ICM | ptx.FLAG_OBSERVE, mode=ptx.MODE_SPLIT)
    a = ptx.Engine(probs, lads, sliced_sweep=waves, **kw)
    b = ptx.Engine(probs, lads, sliced_sweep=-1, **kw)
    a.step(25)
    b.step(25)
    for i in range(2):
        oa, ob = a.observables(i), b.observables(i)
        assert all(np.array_equal(oa[f], ob[f]) for f in oa)
    # a checkpoint of the sliced engine resumed by the per-replica one, and back
    blob = a.save_state()
    b.load_state(blob)
    a.step(15)
    b.step(15)
    _same(_state(a, 2), _state(b, 2))
    b.load_state(a.save_state())
    a.step(5)
    b.step(5)
    _same(_state(a, 2), _state(b, 2))
    a.close()
    b.close()
