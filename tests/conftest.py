import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")


@pytest.fixture(scope="session")
def gpu():
    """Selects device 0 through the C-ABI; the GPU tests fail loudly without the library."""
    import ed_solver_amd as ed
    from ed_solver_amd._lib import check
    check(ed.lib().ed_init(0))
    return ed
