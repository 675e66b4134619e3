import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); runs the HIP library")
    config.addinivalue_line("markers", "slow: long CPU case")


@pytest.fixture(scope="session")
def ctx():
    """One library context on cuda:0 (GPU tests only)."""
    import torch

    from memory_database_amd import olapcore as oc

    assert torch.cuda.is_available(), "GPU test collected on a machine without a GPU"
    c = oc.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ncpu():
    from oracle import pyoracle as po

    return max(1, min(16, po.available_cpus()))
