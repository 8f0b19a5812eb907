import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def orc():
    import oracle as orc_mod
    orc_mod.build()
    return orc_mod


@pytest.fixture(scope="session")
def rpx():
    """The HIP library, initialised on device 0.  Fails (never skips) when the
    library or the device is missing: GPU tests only run where both exist."""
    from rajaperf_amd import abi
    arch = abi.init(0)
    assert arch.startswith("gfx950"), arch
    return abi


@pytest.fixture
def rng():
    return np.random.default_rng(20240607)
