"""Checkpoint and resume of the engine (ptx_engine_save_state / load_state).

An engine stopped after k steps, saved, and loaded into a fresh engine of
the same shape continues bit for bit as an engine that never stopped:
spins, energies, temperature permutation, flips, best energies, sweeps and
states, swap counters, histograms and observables -- also when the resumed
engine runs another execution mode, and with retired instances.  A state of
another shape is refused."""
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


def _batch(ptx):
    probs = [ptx.Problem.read(graph_path("lattice_2d_10.gset")),
             ptx.Problem.from_coo(*instances.ea_3d(6, seed=3)),
             ptx.Problem.from_coo(*instances.erdos_renyi(300, 0.03, seed=2, pm=True))]
    lads = [ptx.beta_linspace(0.5, 3.0, nt) for nt in (6, 8, 5)]
    return probs, lads


def _state(eng, k, observe=False):
    out = []
    for i in range(k):
        be, bs, bst = eng.best(i)
        a, s, n = eng.swap_stats(i)
        row = [eng.spins(i), eng.energies(i), eng.slot2rep(i), eng.flips(i), be, bs, bst, a, s, np.array([n]),
               eng.histograms(i)[0]]
        if observe:
            row += list(eng.observables(i).values())
        out.append(row)
    return out


def _same(x, y):
    for a, b in zip(x, y):
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("save_mode,load_mode", [(1, 1), (2, 2), (3, 3), (1, 3), (3, 2)])
def test_resume_is_bit_exact(ptx, save_mode, load_mode):
    probs, lads = _batch(ptx)
    kw = dict(seed=17, flags=ptx.FLAG_ICM | ptx.FLAG_HISTOGRAM, hist_start=3, steps_per_launch=4)
    say("save mode", save_mode, "load mode", load_mode)
    ref = ptx.Engine(probs, lads, mode=save_mode, **kw)
    ref.step(30)
    a = ptx.Engine(probs, lads, mode=save_mode, **kw)
    a.step(12)
    blob = a.save_state()
    a.step(5)  # the saved engine runs on; the blob is a copy
    b = ptx.Engine(probs, lads, mode=load_mode, **kw)
    b.step(3)  # whatever it did before is overwritten by the load
    b.load_state(blob)
    assert b.info().steps_done == 12
    b.step(18)
    _same(_state(ref, len(probs)), _state(b, len(probs)))
    for e in (ref, a, b):
        e.close()


def test_resume_keeps_retired_instances_and_observables(ptx):
    probs, lads = _batch(ptx)
    kw = dict(seed=5, flags=ptx.FLAG_ICM | ptx.FLAG_OBSERVE | ptx.FLAG_HISTOGRAM, hist_start=2, mode=ptx.MODE_SPLIT)
    ref = ptx.Engine(probs, lads, **kw)
    ref.step(10)
    ref.set_active([True, False, True])
    ref.step(10)
    a = ptx.Engine(probs, lads, **kw)
    a.step(10)
    a.set_active([True, False, True])
    a.step(4)
    blob = a.save_state()
    b = ptx.Engine(probs, lads, **kw)
    b.load_state(blob)
    b.step(6)
    _same(_state(ref, len(probs), observe=True), _state(b, len(probs), observe=True))


def test_state_of_another_shape_is_refused(ptx):
    probs, lads = _batch(ptx)
    a = ptx.Engine(probs, lads, seed=1, flags=ptx.FLAG_ICM)
    a.step(2)
    blob = a.save_state()
    for other in (ptx.Engine(probs, lads, seed=2, flags=ptx.FLAG_ICM),             # another seed
                  ptx.Engine(probs[:2], lads[:2], seed=1, flags=ptx.FLAG_ICM),     # other instances
                  ptx.Engine(probs, [l * 1.01 for l in lads], seed=1, flags=ptx.FLAG_ICM)):  # other ladders
        with pytest.raises(ptx.PtxError) as ex:
            other.load_state(blob)
        assert ex.value.code == -1
    with pytest.raises(ptx.PtxError):
        a.load_state(blob[:100])
