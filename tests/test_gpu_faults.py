"""Fault reporting of the engine, forced from the host so that every kernel
still ends normally: a cooperative step wait that gives up (wait bound of
zero: the first poll that finds a replica missing gives up) must end the
launch, and the next step, sync and read must report PTX_ERR_DEVICE with no
further launch; an engine with the default bound runs the same steps
cleanly.  The histogram fault word stays zero on a correct pilot."""
import numpy as np
import pytest

from conftest import graph_path

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_cooperative_wait_give_up_is_reported(ptx):
    p = ptx.Problem.read(graph_path("lattice_2d_20.gset"))
    betas = ptx.beta_linspace(0.5, 3.0, 16)
    say("cooperative engine, wait bound 0, lattice_2d_20, 16 slots x 2 chains")
    eng = ptx.Engine(p, betas, seed=3, flags=ptx.FLAG_ICM, mode=ptx.MODE_COOP, coop_wait_us=-1)
    eng.step(20)  # enqueued; the first wait gives up inside the launch
    with pytest.raises(ptx.PtxError) as ex:
        eng.sync()
    assert ex.value.code == -4 and "wait gave up" in str(ex.value)
    for read in (lambda: eng.energies(), lambda: eng.spins(), lambda: eng.best(), lambda: eng.min_energies()):
        with pytest.raises(ptx.PtxError) as ex:
            read()
        assert ex.value.code == -4
    with pytest.raises(ptx.PtxError) as ex:
        eng.step(5)  # nothing more is enqueued
    assert ex.value.code == -4
    assert eng.info().steps_done == 20
    eng.close()
    say("default bound: same steps run cleanly")
    ok = ptx.Engine(p, betas, seed=3, flags=ptx.FLAG_ICM, mode=ptx.MODE_COOP)
    ok.step(20)
    ok.sync()
    e = ok.energies()
    s = ok.spins()
    assert all(p.energy(s[r]) == e[r] for r in range(len(e)))
    ok.close()


def test_pilot_histograms_report_no_dropped_samples(ptx):
    p = ptx.Problem.read(graph_path("lattice_2d_20.gset"))
    hb, _, _ = ptx.pilot_betas(0.5, 3.0, 16)
    for mode in (ptx.MODE_SPLIT, ptx.MODE_FUSED, ptx.MODE_COOP):
        say("pilot mode", mode)
        eng = ptx.Engine(p, hb, seed=1, chains=1, flags=ptx.FLAG_PILOT | ptx.FLAG_HISTOGRAM, hist_start=10,
                         mode=mode)
        eng.step(60)
        eng.sync()  # raises if any sample fell outside its table
        counts, _ = eng.histograms()
        assert np.all(counts.sum(axis=1) == 50)
        eng.close()
