import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("photons-2.0_amd", "oracle", os.path.join("tests", "golden")):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


def pytest_addoption(parser):
    parser.addoption("--full", action="store_true", default=False,
                     help="also run the large-size GPU parity cases")


@pytest.fixture(scope="session")
def full_sizes(request):
    return request.config.getoption("--full")
