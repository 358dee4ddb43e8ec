import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ks():
    """The drop-in module, initialised on cuda:0."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from num_amd import kokkos_spmv
    kokkos_spmv.init(0)
    return kokkos_spmv


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
