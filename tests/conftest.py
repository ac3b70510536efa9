import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))

GOLDEN = TESTS / "golden"
REFDATA = GOLDEN / "reference_data"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP libraries")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def _gpu_count() -> int:
    try:
        import nbody_amd
        return nbody_amd.device_count()
    except Exception:
        return 0


_GPUS = None


def pytest_collection_modifyitems(config, items):
    global _GPUS
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    if _GPUS is None:
        _GPUS = _gpu_count()
    if _GPUS == 0:
        skip = pytest.mark.skip(reason="no GPU visible")
        for it in gpu_items:
            it.add_marker(skip)


@pytest.fixture(scope="session", autouse=True)
def _quiet_logs():
    import nbody_amd
    nbody_amd.log_quiet(os.environ.get("NBODY_VERBOSE", "0") == "0")
    yield
