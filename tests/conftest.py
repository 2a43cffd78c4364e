import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the engine through the C ABI")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle

    oracle.load()
    return oracle


def say(*a):
    """Progress line that names the case before a GPU step (flushed)."""
    print(*a, flush=True)


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


# error-model checkers shared with __graft_entry__.smoke() live with the oracle
from oracle.oracle import check_records, fp32c_position_bounds, fp32c_position_error_ok, orbit_sum_bound  # noqa: E402,F401
