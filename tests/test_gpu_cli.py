"""The drop-in command line (`pt_icm`, the reference's flags and result keys,
src/main.cc:33-96 and src/solver.cc:151-163) on the GPU: the printed best
state reproduces the printed best energy and cut on the host."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import REPO, graph_path

pytestmark = pytest.mark.gpu

CLI = os.path.join(REPO, "pt-icm_amd", "build", "pt_icm")


def run(*args):
    out = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:], flush=True)
    assert out.returncode == 0, out.stderr
    keys = dict(re.findall(r"^([a-z_]+): (.*)$", out.stdout, re.M))
    return out.stdout, keys


@pytest.mark.parametrize("args", [
    ("-p", "lattice_2d_10.gset", "-s", "0", "-r", "16", "--pts", "300", "-T0", "0.5", "-T1", "3.0", "--icm"),
    ("-p", "G039", "-s", "3", "-r", "16", "--pts", "200", "-T0", "0.4", "-T1", "2.0", "--preprocess",
     "--preprocess-samples", "200"),
])
def test_cli_prints_a_consistent_best_state(ptx, args):
    args = list(args)
    args[1] = graph_path(args[1])
    text, keys = run(*args)
    for k in ("best_state_energy", "best_state_sweep", "best_energy", "best_maxcut", "num_of_flips",
              "num_of_swaps", "best_state"):
        assert k in keys, k
    assert "Total time:" in text
    state = np.array([int(x) for x in keys["best_state"].strip("[]").split(",")], np.uint8)
    p = ptx.Problem.read(args[1])
    assert len(state) == p.n
    assert float(keys["best_energy"]) == p.energy(state) == float(keys["best_state_energy"])
    assert float(keys["best_maxcut"]) == p.maxcut(state)
    assert int(keys["num_of_flips"]) > 0
