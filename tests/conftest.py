import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def legacy_fixtures():
    return json.loads((GOLDEN / "legacy_fixtures.json").read_text())


@pytest.fixture(scope="session")
def reference_digests():
    return json.loads((GOLDEN / "reference_digests.json").read_text())


@pytest.fixture(scope="session")
def hip():
    """The HIP library; a missing library or device is a test failure for
    gpu-marked tests (there is no CPU fallback to hide behind)."""
    import sufkit_amd

    lib = sufkit_amd.load_hip()
    assert lib.sfk_hip_device_count() >= 1, "no HIP device visible"
    return sufkit_amd
