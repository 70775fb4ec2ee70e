import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(autouse=True)
def _precision_env():
    """Each test chooses its own ALIS_PRECISION; restore the caller's afterwards."""
    saved = os.environ.get("ALIS_PRECISION")
    yield
    if saved is None:
        os.environ.pop("ALIS_PRECISION", None)
    else:
        os.environ["ALIS_PRECISION"] = saved
