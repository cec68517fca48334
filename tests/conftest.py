"""Shared pytest configuration: the `gpu` marker and import paths."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "pennylane-lightning-kokkos_amd"), os.path.join(REPO, "oracle"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)

TOL = float(os.environ.get("TOL", "1e-4"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def capi():
    from pennylane_lightning_kokkos_amd import _capi

    _capi.load()
    if _capi.device_count() < 1:
        pytest.fail("no HIP device visible")
    return _capi


@pytest.fixture
def tol():
    return TOL
