import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "siam-quantum_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
BASIS_DIR = os.path.join(ROOT, "siam-quantum_amd", "data", "basis")
G2_DIR = os.path.join(ROOT, "tests", "golden", "g2")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference_grad():
    """The reference's force routines (src/grad.c) compiled into oracle/_ref (if built)."""
    from oracle import ReferenceGrad
    try:
        return ReferenceGrad()
    except FileNotFoundError:
        pytest.skip("oracle/_ref/libsqref_grad.so not built")


@pytest.fixture(scope="session")
def reference_mp2():
    """The reference's MP2 (src/mp2.c) compiled into oracle/_ref (if built)."""
    from oracle import ReferenceMP2
    try:
        return ReferenceMP2()
    except FileNotFoundError:
        pytest.skip("oracle/_ref/libsqref_mp2.so not built")


@pytest.fixture(scope="session")
def reference():
    """The reference's own hot-path code compiled into oracle/_ref (if built)."""
    from oracle import Reference
    try:
        return Reference()
    except FileNotFoundError:
        pytest.skip("oracle/_ref/libsqref.so not built")
