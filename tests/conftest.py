import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "justpic.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU session (fails loudly without a GPU
    or without the built library)."""
    from justpic_hip import Context
    c = Context(0)
    yield c
    c.close()
