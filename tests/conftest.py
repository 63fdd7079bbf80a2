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
