import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


GOLDEN = ROOT / "tests" / "golden"


@pytest.fixture(scope="session")
def g12():
    import json
    with open(GOLDEN / "g12.json") as f:
        return json.load(f)


@pytest.fixture(scope="session")
def g12_reference():
    import json
    with open(GOLDEN / "g12-reference.json") as f:
        return json.load(f)
