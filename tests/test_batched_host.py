"""Host-side validation of plk_adjoint_batched: every error is raised while
planning, before any device work, so these run without a GPU."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def capi():
    from pennylane_lightning_kokkos_amd import _capi

    _capi.load()
    return _capi


def test_batched_errors(capi):
    ops_ok = capi.Ops(["RX"], [[0.1]], [[0]], [False], [[]])
    z = capi.Obs.named("PauliZ", [0])
    with pytest.raises(capi.PLException, match="circuit 1: .*two wires"):
        capi.adjoint_batched([2, 3], [ops_ok, capi.Ops(["Toffoli", "RX"], [[], [0.2]],
                                                        [[0, 1, 2], [0]], [False] * 2, [[]] * 2)],
                             [z, z])
    with pytest.raises(capi.PLException, match="Pauli words"):
        capi.adjoint_batched([2], [ops_ok], [capi.Obs.hermitian(np.eye(2), [0])])
    with pytest.raises(capi.PLException, match="qubits"):
        capi.adjoint_batched([21], [ops_ok], [z])
    with pytest.raises(capi.PLException, match="No trainable parameters"):
        capi.adjoint_batched([2], [ops_ok], [z], [[]])
    with pytest.raises(capi.PLException, match="not supported using the adjoint"):
        capi.adjoint_batched([2], [capi.Ops(["Rot"], [[0.1, 0.2, 0.3]], [[0]], [False], [[]])],
                             [z], [[0]])
    ev, rows = capi.adjoint_batched([], [], [])
    assert len(ev) == 0 and rows == []


def test_null_state_handles_are_rejected(capi):
    import ctypes

    L = capi.load()
    out = (ctypes.c_double * 4)()
    rc = L.plk_probs(None, None, 0, out)
    assert rc != 0 and "null state handle" in L.plk_last_error().decode()
    rc = L.plk_generate_samples(None, 1, 0, None)
    assert rc != 0
