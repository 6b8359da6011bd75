import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ascend-open-source-contributions_amd")
for p in (ROOT, PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu():
    """A live aclblas handle on cuda:0 (default stream). Skips when no GPU is present."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU in this environment")
    import cgemm3m

    h = cgemm3m.Handle()
    yield h
    h.close()
