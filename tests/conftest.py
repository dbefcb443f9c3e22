import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.lib()
    if os.environ.get("PYTEST_XDIST_WORKER"):
        # several workers share the cores: one OpenMP thread each (spinning
        # thread teams oversubscribed many times over slow the suite 40x)
        O.lib().or_set_num_threads(1)
    return O
