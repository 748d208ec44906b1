"""Shared test configuration: the `gpu` marker and device fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the HIP library")


@pytest.fixture(scope="session")
def dev():
    """cuda:0 with the native library loaded; fails loudly otherwise."""
    import torch
    import apollofft_amd._native as N
    assert torch.cuda.is_available(), "gpu tests need a visible GPU"
    N.lib()  # raises NativeLibraryMissing when the .so was not built
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle
