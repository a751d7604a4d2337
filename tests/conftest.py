import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (runs on MI355X)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        oracle.build(with_ref=False)
    return oracle


@pytest.fixture(scope="session")
def engine_mod():
    """The native engine; GPU tests fail loudly if it is not built."""
    import hiqsimulator_amd
    assert hiqsimulator_amd.device_count() > 0, "no HIP device visible"
    return hiqsimulator_amd
