import os
import sys

import pytest

# Under pytest-xdist every worker would start a full OpenMP team for the oracle's parallel
# loops; on small decks the oversubscribed teams spin against each other (a 2 s oracle run
# then takes minutes). Each worker takes its share of the cores instead. The OpenMP runtime
# reads the variable when it is loaded, so it is set before torch (which may load it).
if "PYTEST_XDIST_WORKER_COUNT" in os.environ:
    _workers = max(1, int(os.environ["PYTEST_XDIST_WORKER_COUNT"]))
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // _workers))

# torch ships its own libamdhip64.so.7, with the same soname as the ROCm runtime the library
# links; a process holds one of them, the first one loaded, and torch's GPU calls work only
# on its own. Tests that take device memory from torch therefore load torch first.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle
