import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import cr_import  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP kernels through the C-ABI")


@pytest.fixture(scope="session")
def pkg():
    return cr_import.package()


@pytest.fixture(scope="session")
def S(pkg):
    return pkg.structs


@pytest.fixture(scope="session")
def cfg(pkg):
    return cr_import.configs()


@pytest.fixture(scope="session")
def O(pkg):
    # CR_ORACLE_LIB: another build of the oracle's sources (the sanitizer build)
    return cr_import.oracle(os.environ.get("CR_ORACLE_LIB"))


@pytest.fixture(scope="session")
def ctx(pkg):
    c = pkg.Context(0)
    yield c
    c.close()
