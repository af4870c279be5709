import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def bfd():
    """The HIP library's Python mirror (package dir learn-zkp_amd/)."""
    from __graft_entry__ import load_package
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    import os
    import oracle as O
    O.lib()
    if not os.environ.get("BFO_THREADS"):
        O.set_threads(min(16, os.cpu_count() or 1))  # exact arithmetic: results do not depend on it
    return O
