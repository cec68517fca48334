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


def test_batch_eligibility():
    """Tapes the batched engine takes: no state preparation, no explicit
    matrix, operations on at most two wires, one Pauli-sum observable."""
    from pennylane_lightning_kokkos_amd.device import batch_eligible

    ham = ("ham", [0.5, -1.0], [("named", "PauliZ", [0]),
                                ("tensor", [("named", "PauliX", [1]), ("named", "PauliY", [2])])])
    base = [Op("Rot", [0], [0.1, 0.2, 0.3]), Op("CNOT", [0, 1]), Op("IsingXX", [1, 2], [0.4])]

    def ok(ops, obs):
        ser, sp = serialize_ops(Tape(ops, obs))
        return batch_eligible(ser, sp, obs)

    assert ok(base, [ham])
    assert ok(base, [("named", "PauliZ", [2])])
    assert not ok(base, [ham, ("named", "PauliZ", [2])])                  # two observables
    assert not ok(base, [("herm", np.eye(2), [0])])                       # not a Pauli sum
    assert not ok(base + [Op("Toffoli", [0, 1, 2])], [ham])               # three wires
    assert not ok([Op("BasisState", [0], [np.array([1])])] + base, [ham])  # state preparation
    u = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    assert not ok(base + [Op("QubitUnitary", [1], [], False, u)], [ham])  # explicit matrix
