import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device (MI355X)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def oracle_release():
    """The same restatement built with the reference's Release flags (-O3 -ffast-math,
    CMakeLists.txt:71): how far the reference as shipped departs from strict IEEE."""
    from oracle.oracle import Oracle

    return Oracle(fast=True)


@pytest.fixture(scope="session")
def reference():
    from oracle.oracle import Reference

    if not Reference.available():
        pytest.skip("compiled reference (oracle/_ref) not present on this machine")
    return Reference()


@pytest.fixture(scope="session")
def engine():
    import granite_amd

    return granite_amd.Engine(0)
