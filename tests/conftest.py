import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "cbm-gpu_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")


@pytest.fixture(autouse=True, scope="session")
def _compressed_form_only():
    """The suite checks the CBM kernels themselves: matrices are built
    without the uncompressed form unless a test asks for it
    (tests/test_gpu_uncompressed.py).  Session-wide, so that module-scoped
    matrices are built the same way."""
    try:
        from cbm import cbm4mm
    except Exception:  # the package does not import (no torch): nothing to switch
        yield
        return
    saved = cbm4mm.UNCOMPRESSED
    cbm4mm.UNCOMPRESSED = False
    yield
    cbm4mm.UNCOMPRESSED = saved
