"""The pybind11 module `lightning_kokkos_qubit_ops` on the GPU, written the way
the reference's Python tests use the module (tests/test_adjoint_jacobian.py,
test_hamiltonian.py, test_hamiltonian_sparse.py, test_vjp.py): same class
names, same call signatures, closed forms and goldens of those tests."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import dense_pauli_sparse, random_hamiltonian, random_ops, random_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    from pennylane_lightning_kokkos_amd import lightning_kokkos_qubit_ops as mod

    return mod


def _sv(m, n, dtype=np.complex128):
    return (m.LightningKokkos_C128 if dtype == np.complex128 else m.LightningKokkos_C64)(n)


def _adj(m, dtype, sv, obs, names, params, wires, inverses, tp):
    sfx = "C128" if dtype == np.complex128 else "C64"
    adj = getattr(m, f"AdjointJacobianKokkos_{sfx}")()
    rdt = np.float64 if dtype == np.complex128 else np.float32
    ops = adj.create_ops_list(names, [np.array(p, dtype=rdt) for p in params], wires, inverses,
                              [np.array([], dtype=dtype) for _ in names])
    return np.array(adj.adjoint_jacobian(sv, obs, ops, tp))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("par", [1, -2, 1.623, -0.051, 0])
def test_ry_gradient(m, dtype, par):
    """test_adjoint_jacobian.py:226-244: d<X>/d(par) of RY(par)|0> = cos(par)."""
    sv = _sv(m, 1, dtype)
    sv.RY([0], False, [par])
    sfx = "C128" if dtype == np.complex128 else "C64"
    obs = [getattr(m, f"NamedObsKokkos_{sfx}")("PauliX", [0])]
    jac = _adj(m, dtype, sv, obs, ["RY"], [[par]], [[0]], [False], [0])
    assert jac[0, 0] == pytest.approx(np.cos(par), abs=1e-6 if dtype == np.complex64 else 1e-12)


def test_multiple_rx_gradient(m):
    """test_adjoint_jacobian.py:259-278: -diag(sin(params))."""
    params = np.array([np.pi, np.pi / 2, np.pi / 3])
    sv = _sv(m, 3)
    for w in range(3):
        sv.RX([w], False, [params[w]])
    obs = [m.NamedObsKokkos_C128("PauliZ", [w]) for w in range(3)]
    jac = _adj(m, np.complex128, sv, obs, ["RX"] * 3, [[p] for p in params], [[0], [1], [2]],
               [False] * 3, [0, 1, 2])
    np.testing.assert_allclose(jac, -np.diag(np.sin(params)), atol=1e-12)


@pytest.mark.parametrize("obs, coeffs, res", [
    ([("PauliX", 0), ("PauliZ", 1)], None, 0.0),
    ([("PauliZ", 0), ("PauliZ", 1)], None, math.cos(0.4) * math.cos(-0.2)),
])
def test_expval_tensor_after_rotations(m, obs, coeffs, res):
    """test_hamiltonian.py:29-52 with the tensor observable."""
    sv = _sv(m, 2)
    sv.RX([0], False, [0.4])
    sv.RY([1], False, [-0.2])
    names = [o[0] for o in obs]
    wires = [o[1] for o in obs]
    pauli = {"PauliX": np.array([[0, 1], [1, 0]]), "PauliZ": np.diag([1, -1])}
    mat = np.kron(pauli[names[0]], pauli[names[1]]).astype(np.complex128)
    assert sv.ExpectationValue(names, wires, [], mat.ravel()) == pytest.approx(res, abs=1e-12)
    t = m.TensorProdObsKokkos_C128([m.NamedObsKokkos_C128(n, [w]) for n, w in obs])
    ham = m.HamiltonianKokkos_C128(np.array([1.0, 0.2]), [t, m.NamedObsKokkos_C128("PauliZ", [0])])
    assert "PauliZ[0]" in repr(ham) and t.get_wires() == [0, 1]


def test_hamiltonian_expectation_state_vector(m):
    """test_hamiltonian.py:54-81: H = 3.1415 I + 9.6 X1 Y2 on a fixed state; expected 3.1415."""
    st = np.array([0.0, 0.1j, 0.1 + 0.1j, 0.1 + 0.2j, 0.2 + 0.2j, 0.3 + 0.3j, 0.3 + 0.4j, 0.4 + 0.5j],
                  dtype=np.complex128)
    sv = _sv(m, 3)
    sv.HostToDevice(st)
    d, i, p = dense_pauli_sparse(3, ("ham", [3.1415, 9.6], [("named", "Identity", [1]),
                                                           ("tensor", [("named", "PauliX", [1]),
                                                                       ("named", "PauliY", [2])])]))
    assert sv.ExpectationValue(d, list(i), list(p)) == pytest.approx(3.1415, rel=1e-12)


@pytest.mark.parametrize("word, expected", [
    ((("PauliX", 0),), 0.0), ((("PauliX", 1),), -0.19866933079506122),
    ((("PauliY", 0),), -0.38941834230865050), ((("PauliY", 1),), 0.0),
    ((("PauliZ", 0),), 0.92106099400288520), ((("PauliZ", 1),), 0.98006657784124170)])
def test_sparse_pauli_words(m, word, expected):
    """test_hamiltonian_sparse.py:58-79: sparse Pauli words after RX(0.4), RY(-0.2)."""
    sv = _sv(m, 2)
    sv.RX([0], False, [0.4])
    sv.RY([1], False, [-0.2])
    ob = ("named", word[0][0], [word[0][1]])
    d, i, p = dense_pauli_sparse(2, ("ham", [1.0], [ob]))
    assert sv.ExpectationValue(d, list(i), list(p)) == pytest.approx(expected, abs=1e-12)
    sh = m.SparseHamiltonianKokkos_C128(d, list(i), list(p), [0, 1])
    assert sh.get_wires() == [0, 1]


def test_vjp_equals_adjoint_of_weighted_hamiltonian(m):
    """test_vjp.py:41-71: vjp with dy == adjoint Jacobian of sum_i dy_i O_i."""
    rng = np.random.default_rng(3)
    n = 4
    names, params, wires, inv, _ = random_ops(n, 25, rng, names=["RX", "RY", "RZ", "CNOT", "IsingZZ"])
    npar = sum(1 for q in params if q)
    sv = _sv(m, n)
    for nm, p, w, b in zip(names, params, wires, inv):
        getattr(sv, nm)(w, b, p)
    obs = [m.NamedObsKokkos_C128("PauliZ", [0]), m.NamedObsKokkos_C128("PauliX", [2])]
    jac = _adj(m, np.complex128, sv, obs, names, params, wires, inv, list(range(npar)))
    dy = np.array([0.7, -1.3])
    ham = m.HamiltonianKokkos_C128(dy, obs)
    vjp = _adj(m, np.complex128, sv, [ham], names, params, wires, inv, list(range(npar)))
    np.testing.assert_allclose(vjp[0], dy @ jac, atol=1e-12)


def test_module_state_methods_match_oracle(m):
    rng = np.random.default_rng(9)
    n = 6
    psi = random_state(n, rng)
    sv = m.LightningKokkos_C128(psi)
    assert sv.numQubits() == n and sv.dataLength() == 2**n
    sv.apply(["RX", "CNOT", "IsingXY"], [[0], [1, 4], [2, 5]], [False, True, False], [[0.3], [], [1.1]])
    ref = psi.copy()
    for nm, w, b, p in (("RX", [0], False, [0.3]), ("CNOT", [1, 4], True, []), ("IsingXY", [2, 5], False, [1.1])):
        O.apply(ref, nm, w, b, p)
    out = np.zeros(2**n, dtype=np.complex128)
    sv.DeviceToHost(out)
    np.testing.assert_allclose(out, ref, atol=1e-12)
    U = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
    sv.apply("QubitUnitary", [3, 1], False, [], U.ravel())
    O.apply_matrix(ref, U, [3, 1])
    sv.DeviceToHost(out)
    np.testing.assert_allclose(out, ref, atol=1e-12)
    s = sv.applyGenerator("IsingXX", [0, 2], False, [])
    assert s == -0.5
    sv.resetKokkos()
    sv.DeviceToHost(out)
    assert out[0] == 1
    sv.setBasisState(5)
    sv.DeviceToHost(out)
    assert out[5] == 1
    with pytest.raises(m.PLException, match="Generator does not exist"):
        sv.applyGenerator("CNOT", [0, 1], False, [])
