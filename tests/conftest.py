import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libtnx.so")


@pytest.fixture(scope="session")
def fixtures():
    return json.loads((ROOT / "tests" / "golden" / "fixtures.json").read_text())


@pytest.fixture(scope="session")
def gpu():
    """The GPU device; the library must be built (no CPU fallback exists)."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu-marked test run without a GPU")
    from tnpy_amd import _lib

    _lib.load_library()
    return torch.device("cuda", 0)
