import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cedarsim.jl_amd"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcedarhip.so on cuda:0)")


@pytest.fixture(scope="session")
def engine():
    """The HIP engine.  GPU tests fail loudly (no skip) when no device is visible."""
    import cedarhip

    print(f"[gpu] creating engine on device 0", flush=True)
    eng = cedarhip.Engine(0)
    yield eng
    eng.close()
