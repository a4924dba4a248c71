import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU cases")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    try:
        O.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    # the oracle's OpenMP loops cost more than they save on the small grids of the
    # CPU checks (and oversubscribe the host under parallel test workers); tests that
    # step large grids set their own thread count
    O.set_threads(1)
    return O
