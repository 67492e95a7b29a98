import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libmlbvh.so")
    config.addinivalue_line("markers", "slow: large CPU oracle runs")


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle

    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """The device-side API; fails loudly (no fallback) when the library or the
    GPU is missing."""
    import torch

    import minimumlbvh_amd
    from minimumlbvh_amd import lbvh

    minimumlbvh_amd.load()
    assert torch.cuda.is_available(), "gpu tests need a ROCm GPU"
    lbvh.context(0)
    return lbvh


@pytest.fixture
def dev():
    import torch

    return torch.device("cuda:0")


def to_dev(a, dev, dtype=None):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev)
