import os
import sys

import pytest

# torch's ROCm wheel carries its own HIP runtime under the same soname as the system one.
# Imported first, it is the runtime libeio_hip.so binds to, so one process can run the C ABI
# and the torch FWI layer; loaded after libeio_hip.so, torch finds no GPU.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle
