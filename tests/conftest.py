"""Shared test set-up: import paths, the `gpu` marker, common fixtures."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "simsdp-generic-imaging-pipeline_amd")
ORACLE = os.path.join(REPO, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import sdp_oracle

    sdp_oracle.load()
    return sdp_oracle


@pytest.fixture(scope="session")
def hip():
    """The HIP library; GPU tests fail (never skip) when it cannot run."""
    import sdp_hip

    lib = sdp_hip.load()
    sdp_hip.check(lib.sdp_init())
    return sdp_hip


@pytest.fixture(scope="session")
def actors(hip):
    return hip.Actors()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
