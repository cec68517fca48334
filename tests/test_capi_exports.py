"""The C ABI library loads and exports every entry point of include/plk_amd.h
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "plk_amd.h")
LIB = os.path.join(REPO, "pennylane-lightning-kokkos_amd", "pennylane_lightning_kokkos_amd", "libplk_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(plk_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("plk_state_create", "plk_apply", "plk_apply_generator", "plk_expval_obs",
                 "plk_adjoint_jacobian", "plk_obs_hamiltonian", "plk_ops_create", "plk_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built (run python __graft_entry__.py)")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (plk_[a-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_ctypes_front_end_covers_header():
    from pennylane_lightning_kokkos_amd import _capi

    assert set(_capi.SIGNATURES) == set(declared())
    assert _capi.load() is not None


def test_last_error_is_empty_string_initially():
    from pennylane_lightning_kokkos_amd import _capi

    assert isinstance(_capi.load().plk_last_error(), bytes)
