"""Host side of the dense-Hamiltonian route: the dense matrix the device
builds for a Hamiltonian equals the sum of the terms' matrices (numpy kron
in the wire order of first appearance), and the route is taken only for
Hamiltonians that are not Pauli-word sums."""
import numpy as np

from pennylane_lightning_kokkos_amd import device as D


def test_obs_matrix_matches_kron():
    rng = np.random.default_rng(1)
    H = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    H = H + H.conj().T
    X, Z = D._NAMED["PauliX"], D._NAMED["PauliZ"]
    desc = ("ham", [0.5, -1.2, 0.3],
            [("herm", H, [2, 0]), ("tensor", [("named", "PauliX", [1]), ("named", "PauliZ", [0])]),
             ("named", "PauliZ", [2])])
    W = D._obs_wires(desc)
    assert W == [2, 0, 1]
    # Wire order [2, 0, 1]: H already acts on (2, 0), the first two wires.
    ref = 0.5 * np.kron(H, np.eye(2)) - 1.2 * np.kron(np.kron(np.eye(2), Z), X) \
        + 0.3 * np.kron(Z, np.eye(4))
    np.testing.assert_allclose(D._obs_matrix(desc, W), ref, atol=1e-14)


def test_embed_permutes_wires():
    rng = np.random.default_rng(2)
    U = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    # U on (b, a) within [a, b] is SWAP U SWAP.
    S = np.eye(4)[[0, 2, 1, 3]]
    np.testing.assert_allclose(D._embed(U, ["b", "a"], ["a", "b"]), S @ U @ S, atol=1e-14)


def test_route_selection():
    assert D._pauli_only(("ham", [1.0], [("named", "PauliZ", [0])]))
    assert not D._pauli_only(("ham", [1.0], [("herm", np.eye(2), [0])]))
