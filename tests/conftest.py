import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# every compaction the tests run also checks its output in full (offsets,
# entries, row order, and the merged sorted row copy entry by entry)
os.environ.setdefault("MISFIT_COMPACT_SELFCHECK", "1")

GOLDEN = os.path.join(ROOT, "tests", "golden")
SAMPLE = os.path.join(GOLDEN, "sample")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle._LIB):
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def engine_mod():
    import misfit_amd
    misfit_amd.lib()  # raises when the HIP library is not built
    return misfit_amd
