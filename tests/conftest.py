import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def gpu():
    """Skips when no GPU is visible; the GPU suite runs on an MI355X."""
    import torch  # noqa: F401  (device probe only)
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return 0
