import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")
    config.addinivalue_line(
        "markers", "ref: needs oracle/_ref (the reference compiled from source)")


def _ensure_built():
    """Build only when nothing is built yet (after a source change, run
    __graft_entry__.build() first)."""
    from vpic_amd import _lib
    oracle_so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(oracle_so) or not os.path.exists(_lib.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()


_ensure_built()


@pytest.fixture(scope="session")
def orc():
    from oracle.oracle import Oracle
    return Oracle.get()


@pytest.fixture(scope="session")
def ref():
    from oracle.oracle import Ref
    if not Ref.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return Ref.get()
