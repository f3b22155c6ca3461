import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); run with -m gpu")
    config.addinivalue_line("markers", "slow: long CPU test")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def bath_of(npz, prefix="bath_", suffix=""):
    return tuple(npz[f"{prefix}{k}{suffix}"] for k in ("etal", "etar", "etaa", "expn", "mode"))


def pytest_collection_modifyitems(config, items):
    # print the case name before each GPU test so a crash names it
    for it in items:
        if it.get_closest_marker("gpu"):
            it.add_marker(pytest.mark.usefixtures("_announce"))


@pytest.fixture
def _announce(request):
    print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
