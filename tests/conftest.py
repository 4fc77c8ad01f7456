import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def expected():
    import json
    with open(os.path.join(GOLDEN, "expected.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def gpu():
    """The HIP engine on cuda:0.  Fails (does not skip) when no device is visible:
    a GPU test must never pass on a fallback."""
    import dyn_maxflow_amd as dmf
    n = dmf.device_count()
    assert n > 0, "no HIP device visible: GPU tests need an MI355X"
    return dmf
