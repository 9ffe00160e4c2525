import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libsundog.so")


@pytest.fixture(scope="session")
def lib():
    import sundog_amd
    return sundog_amd.load()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def asset_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("assets"))
