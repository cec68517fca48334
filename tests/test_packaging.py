"""Packaging metadata: the pennylane.plugins entry point of the reference
(setup.py:132-136) resolves to this package's device class without
PennyLane."""
import importlib
import os

import tomli

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PYPROJECT = os.path.join(REPO, "pennylane-lightning-kokkos_amd", "pyproject.toml")


def test_plugin_entry_point_resolves():
    meta = tomli.load(open(PYPROJECT, "rb"))
    eps = meta["project"]["entry-points"]["pennylane.plugins"]
    assert set(eps) == {"lightning.kokkos"}
    module, attr = eps["lightning.kokkos"].split(":")
    cls = getattr(importlib.import_module(module), attr)
    assert cls.short_name == "lightning.kokkos"
    assert callable(getattr(cls, "adjoint_jacobian"))


def test_package_data_ships_native_libraries():
    meta = tomli.load(open(PYPROJECT, "rb"))
    assert "*.so" in meta["tool"]["setuptools"]["package-data"]["pennylane_lightning_kokkos_amd"]
