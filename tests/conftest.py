import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_count() -> int:
    try:
        import torch  # noqa: F401  (device counting does not initialise HIP)
        return torch.cuda.device_count()
    except Exception:
        return 0


@pytest.fixture(scope="session")
def ctx():
    from luz_amd import api
    c = api.Context(0)
    yield c
    c.close()
