import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "wwocl-boinc_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def gctx():
    """One device context for the GPU tests, sized for 2^32-wide chunks anywhere in the
    64-bit domain."""
    import wwlib

    cap = wwlib.list_bound(1 << 32, (1 << 32) + (1 << 32))
    with wwlib.Context(0, list_capacity=cap) as ctx:
        yield ctx
