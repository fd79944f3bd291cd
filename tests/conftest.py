import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(autouse=True)
def _exact_oracle_energies(request):
    """GPU parity tests compare energies with the oracle's 80-bit sums: the
    GPU's fixed-order tree sums (and its 80-bit self energy) are closer to
    the exact value than the reference's FP64 running sum, which on large
    boxes is off by ~1e-11 relative (tests/energy_tol.py)."""
    if request.node.get_closest_marker("gpu") is None:
        yield
        return
    import oracle
    oracle.build()
    with oracle.exact_energy():
        yield
