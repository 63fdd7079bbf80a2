"""The C++ facade as a reference caller sees it (pt-icm_amd/csrc/cli/pticm.hh):
tests/facade_caller.cpp is written against the reference's interface
(Problem, Replica, beta_linspace; include/problem.hh, include/replica.hh) and
built against the facade.  Here, without a GPU: the incremental dE update
(src/problem.cc:164-181) equals a full recompute after every single flip,
energy of bool arrays, the host random stream.  The GPU half (Replica::sweep)
runs in tests/test_gpu_cli.py."""
import os
import subprocess

import pytest

from conftest import REPO, graph_path

CALLER = os.path.join(REPO, "pt-icm_amd", "build", "facade_caller")


@pytest.mark.parametrize("name", ["lattice_2d_5.gset", "lattice_2d_10.gset", "G039"])
def test_reference_style_caller_host_checks(name):
    assert os.path.exists(CALLER), "build the library first (make -C pt-icm_amd)"
    out = subprocess.run([CALLER, "host", graph_path(name)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
