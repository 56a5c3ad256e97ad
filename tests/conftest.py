import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")


@pytest.fixture
def double():
    import xgrid_amd
    xgrid_amd.init(precision="double", cacheroot=os.path.join(ROOT, ".xgrid"))
    return xgrid_amd


@pytest.fixture
def single():
    import xgrid_amd
    xgrid_amd.init(precision="float", cacheroot=os.path.join(ROOT, ".xgrid"))
    return xgrid_amd
