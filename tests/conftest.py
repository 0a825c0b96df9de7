"""Shared test configuration.

Markers: ``gpu`` tests need an MI355X (they call the HIP library through
its C ABI); everything else runs on the CPU.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")


def gpu_present() -> bool:
    import siting_amd
    try:
        return siting_amd.device_count() > 0
    except siting_amd.SitingError:
        return False


@pytest.fixture(scope="session")
def sa():
    import siting_amd
    siting_amd.lib()  # fails loudly when the HIP library is not built
    return siting_amd


@pytest.fixture(scope="session")
def gpu(sa):
    if sa.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return sa


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def make_dem(sa, n, rough_h=0.8, hmax=20000, seed=1, nhills=8, sigma=80.0):
    return sa.dem_generate(sa.TerrainSpec(nrows=n, rough_h=rough_h, hmin=0, hmax=hmax,
                                          nhills=nhills, hill_peak=5000.0,
                                          hill_sigma=sigma, seed=seed))
