import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Large anelastic blocks search for a fast arena placement at creation (up to
# six allocations, seconds each); the tests take the first arena, except
# tests/test_gpu_arena.py, which runs the search itself.
os.environ.setdefault("WQL3D_ARENA_TRIES", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O
