"""The pybind11 module exposes the reference module's names
(pennylane_lightning_kokkos/src/bindings/Bindings.cpp:59-968); callers
dispatch on these names with getattr (_serialize.py:294, lightning_kokkos.py:378)."""
import pytest

GATES = ["Identity", "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "CNOT", "SWAP", "CSWAP",
         "Toffoli", "CY", "CZ", "PhaseShift", "ControlledPhaseShift", "RX", "RY", "RZ", "Rot", "CRX",
         "CRY", "CRZ", "CRot", "IsingXX", "IsingXY", "IsingYY", "IsingZZ", "MultiRZ",
         "SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus", "DoubleExcitation",
         "DoubleExcitationMinus", "DoubleExcitationPlus"]
STATE_METHODS = ["setBasisState", "setStateVector", "apply", "applyGenerator", "ExpectationValue",
                 "DeviceToHost", "HostToDevice", "numQubits", "dataLength", "resetKokkos"]


@pytest.fixture(scope="module")
def ops_mod():
    from pennylane_lightning_kokkos_amd import lightning_kokkos_qubit_ops as m

    return m


@pytest.mark.parametrize("sfx", ["C64", "C128"])
def test_class_names(ops_mod, sfx):
    for cls in ("LightningKokkos", "ObservableKokkos", "NamedObsKokkos", "HermitianObsKokkos",
                "TensorProdObsKokkos", "HamiltonianKokkos", "SparseHamiltonianKokkos",
                "OpsStructKokkos", "AdjointJacobianKokkos"):
        assert hasattr(ops_mod, f"{cls}_{sfx}"), cls


@pytest.mark.parametrize("sfx", ["C64", "C128"])
def test_state_methods(ops_mod, sfx):
    sv = getattr(ops_mod, f"LightningKokkos_{sfx}")
    for name in GATES + STATE_METHODS:
        assert hasattr(sv, name), name
    # Operations without a kernel are absent, so the serializer falls back to matrices.
    assert not hasattr(sv, "QubitUnitary") and not hasattr(sv, "ISWAP")


def test_module_functions(ops_mod):
    for f in ("kokkos_start", "kokkos_end", "kokkos_config_info", "print_configuration",
              "InitializationSettings", "PLException"):
        assert hasattr(ops_mod, f)
    s = ops_mod.InitializationSettings()
    assert s.get_device_id() == 0 and not s.has_device_id()
    s.set_device_id(0).set_num_threads(2)
    assert s.has_device_id() and s.get_num_threads() == 2
    assert "device_id = 0" in repr(s)
    assert hasattr(ops_mod.AdjointJacobianKokkos_C128(), "create_ops_list")
