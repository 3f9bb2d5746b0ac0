import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def oracle_ms_cache():
    from oracle.pyoracle import OracleModelSpace
    cache = {}

    def get(emax, A=16, Z=8):
        key = (emax, A, Z)
        if key not in cache:
            cache[key] = OracleModelSpace(emax, A, Z)
        return cache[key]

    return get
