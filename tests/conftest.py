import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch bundles its own HIP runtime under the same soname as ROCm's.  Loading
# torch first makes libbmmpy_hip.so bind to that one runtime; loading our
# library first would put two HIP runtimes in the process and torch would then
# see no device.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")


@pytest.fixture(scope="session")
def ora():
    import oracle
    if not os.path.exists(oracle.ORACLE_SO):
        oracle.build(ref=False)
    return oracle.oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle
    r = oracle.reference()
    if r is None:
        pytest.skip("reference CPU library not built here")
    return r


@pytest.fixture(scope="session")
def lib():
    from bmmpy_amd import _lib
    return _lib
