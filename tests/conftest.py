import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def dev():
    """cuda:0 for GPU tests. Fails loudly (no skip) if no GPU is visible."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    from phix_amd import capi
    capi.lib()
    return torch.device("cuda:0")
