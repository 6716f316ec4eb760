import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library's compute paths)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def ctx():
    import solidfem_amd as sf
    return sf.Context(0)


@pytest.fixture
def forced_variant():
    """Forces a tuning variant through the library (no setenv while library threads run); undone after the test."""
    import solidfem_amd as sf
    yield sf.force_variant
    sf.force_variant(0)


@pytest.fixture
def slot_limit():
    """Lowers the device setup's per-row block limit through the library; undone after the test."""
    import solidfem_amd as sf
    yield sf.setup_slot_limit
    sf.setup_slot_limit(0)
