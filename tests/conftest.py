import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def oracle_threads() -> int:
    """Threads for the OpenMP oracle loops (the GPU box grants 16 cores)."""
    return max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="session")
def ctx():
    import seismic_code_github_amd as sfd
    c = sfd.Context(0)
    yield c
    c.close()
