import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run with -m gpu")


@pytest.fixture(scope="session")
def ref_available():
    from oracle import REF_DIR

    return os.path.exists(os.path.join(REF_DIR, "ref_cpu6.so"))


@pytest.fixture(scope="session", autouse=True)
def _built():
    """Build the native artefacts if this checkout has not been built yet."""
    import hermite_amd
    import oracle

    if not (os.path.exists(hermite_amd.HIP_LIB) and os.path.exists(hermite_amd.HOST_LIB)):
        hermite_amd.build()
    if not os.path.exists(os.path.join(oracle.HERE, "liboracle.so")):
        oracle.build()
