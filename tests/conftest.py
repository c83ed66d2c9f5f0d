import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


def load_nbg():
    """The product package (directory n-body_amd/, imported by path)."""
    return importlib.import_module("n-body_amd")


@pytest.fixture(scope="session")
def nbg():
    return load_nbg()


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx(nbg):
    """A context on device 0.  The library must load; a machine without a GPU
    skips, a GPU machine where the library fails raises."""
    nbg.lib()
    if nbg.device_count() == 0:
        pytest.skip("no GPU visible")
    ctx = nbg.Context(0)
    yield ctx
    ctx.close()
