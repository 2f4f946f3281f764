import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libkete_hip.so")


@pytest.fixture(scope="session")
def spk_path():
    from kete_amd.synthetic import ephemeris_path
    return ephemeris_path()


@pytest.fixture(scope="session")
def oracle_spk(spk_path):
    import oracle
    assert oracle.spk_load(spk_path) == 0
    return oracle


@pytest.fixture(scope="session")
def ctx(spk_path):
    import kete_amd
    return kete_amd.get_context()
