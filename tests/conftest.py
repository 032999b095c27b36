import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU-oracle runs")


@pytest.fixture(scope="session")
def luts():
    from sunaengine_amd import tables
    return tables.load_luts()


@pytest.fixture(scope="session")
def mats_img():
    from sunaengine_amd import materials
    return materials.mats_image()
