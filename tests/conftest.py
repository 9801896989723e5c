"""pytest configuration: the `gpu` marker and shared fixtures.

CPU tests (-m "not gpu") check the oracle against closed forms, exact
arithmetic and a third-party LAPACK, the host logic, and that the C-ABI
library loads and exports every symbol the headers declare.  GPU tests (-m gpu)
run the library's kernels on the device and compare with the oracle.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def queue():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import magma_amd
    q = magma_amd.Queue(0)
    yield q
    q.close()
