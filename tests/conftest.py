import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def refpath():
    from oracle.oracle import LIB_REF, RefPath
    if not os.path.exists(LIB_REF):
        pytest.skip("oracle/_ref/libref_path.so not built (needs the reference sources)")
    return RefPath()


@pytest.fixture(scope="session")
def capi():
    """the ctypes binding of the HIP library (loading fails loudly when it is not built)"""
    from rpmd_mainz_amd import capi as c
    c.lib()
    return c
