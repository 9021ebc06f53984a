"""pytest configuration: the ``gpu`` marker and shared fixtures.

GPU tests compare the HIP path (through the C ABI) with the CPU oracle; CPU
tests pin the oracle against the reference's own known answers and check the
library's host side.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def orc():
    import oracle as O

    if not os.path.exists(O.LIB_PATH):
        O.build()
    O.lib()
    return O


@pytest.fixture(scope="session")
def ref_leaves(orc):
    """The flang-built reference leaves, or skip when they are absent."""
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref/libadsref.so not built (needs /root/reference and flang)")
    return R
