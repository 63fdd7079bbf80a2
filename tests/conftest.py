import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests"))
REF_GRAPHS = os.path.join(REPO, "tests", "golden", "graphs")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libptx.so")
    config.addinivalue_line("markers", "slow: long-running statistical check")
    config.addinivalue_line("markers", "gpu_subprocess: runs the GPU in a child process (pt_icm, facade_caller)")


def pytest_collection_modifyitems(config, items):
    """GPU tests that run the engine in a child process go first.  Once a
    process has made a cooperative launch, the HIP runtime keeps the device's
    cooperative-launch resource for that process until it exits, and a
    cooperative launch from another process then fails ("the operation cannot
    be performed in the present state"); the pytest process itself makes
    cooperative launches in many tests.  The order of the other tests is kept."""
    first = [it for it in items if it.get_closest_marker("gpu_subprocess")]
    rest = [it for it in items if not it.get_closest_marker("gpu_subprocess")]
    items[:] = first + rest


def load_ptx():
    """Imports pt-icm_amd/ptx.py (the package directory name is not a Python identifier)."""
    if "ptx" in sys.modules:
        return sys.modules["ptx"]
    path = os.path.join(REPO, "pt-icm_amd", "ptx.py")
    spec = importlib.util.spec_from_file_location("ptx", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ptx"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ptx():
    return load_ptx()


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib


def graph_path(name):
    return os.path.join(REF_GRAPHS, name)
