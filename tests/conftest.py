import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "spatial-query_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def ctx():
    """One engine context on cuda:0 for the whole GPU session."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import sq_amd
    c = sq_amd.Context(0)
    yield c
    c.close()
