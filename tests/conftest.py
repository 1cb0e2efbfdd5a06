import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "gadget-ng_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP product path)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def pm():
    """The product package; GPU tests fail (not skip) when no device is usable."""
    import gadget_ng_pm_hip as pkg
    L = pkg.lib()
    if L.gng_pm_device_count() <= 0:
        pytest.fail("no HIP device visible: GPU tests must run on the GPU box")
    return pkg


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
