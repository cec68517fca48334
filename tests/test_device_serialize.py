"""The device serializer (reference _serialize.py:252-308) without a GPU."""
import numpy as np

from pennylane_lightning_kokkos_amd.device import LightningKokkosAMD, Tape, serialize_ops
from pennylane_lightning_kokkos_amd.torch_adjoint import Op


def test_rot_expands_and_stateprep_is_skipped():
    tape = Tape([Op("StatePrep", [0, 1], [np.array([1, 0, 0, 0])]),
                 Op("Rot", [1], [0.1, 0.2, 0.3]), Op("CNOT", [0, 1])], [("named", "PauliZ", [0])])
    (names, params, wires, inv, mats), sp = serialize_ops(tape)
    assert sp
    assert names == ["RZ", "RY", "RZ", "CNOT"]
    assert params == [[0.1], [0.2], [0.3], []]
    assert wires == [[1], [1], [1], [0, 1]] and inv == [False] * 4


def test_matrix_ops_travel_inverted_with_flag_cleared():
    u = np.array([[0, 1j], [1, 0]], dtype=np.complex128)
    (names, params, wires, inv, mats), sp = serialize_ops(
        Tape([Op("QubitUnitary", [2], [], True, u), Op("RX", [0], [0.5], True)], []))
    assert not sp and names == ["QubitUnitary", "RX"]
    np.testing.assert_array_equal(mats[0], u.conj().T.reshape(-1))
    assert inv == [False, True] and params == [[], [0.5]] and list(mats[1]) == []


def test_trainable_bookkeeping_counts_stateprep_parameter():
    tape = Tape([Op("StatePrep", [0], [np.array([1, 0])]), Op("RX", [0], [0.3]),
                 Op("RY", [0], [0.4])], [], trainable_params=[2, 0, 1])
    assert tape.get_trainable() == [0, 1, 2]
    assert [o for o, _ in tape.all_params()] == [0, 1, 2]
    assert LightningKokkosAMD.short_name == "lightning.kokkos"
