import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU run")


@pytest.fixture(scope="session")
def oracle():
    from oracle_bind import load_oracle, ORACLE_SO
    if not os.path.exists(ORACLE_SO):
        pytest.fail("oracle/liboracle.so is not built (python taylor_amd/build.py)")
    return load_oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle_bind import have_ref, load_ref
    if not have_ref():
        pytest.skip("reference build (oracle/_ref) not present")
    return load_ref()
