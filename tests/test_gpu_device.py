"""The lightning.kokkos-style device on the GPU, written like the reference's
device tests (tests/test_adjoint_jacobian.py, test_vjp.py, test_probs.py,
test_sample.py): closed forms, goldens, the StatePrep column rule, vjp ==
dy @ J, and the device-level errors."""
import numpy as np
import pytest

import oracle as O
from helpers import to_oracle
from pennylane_lightning_kokkos_amd.device import (DeviceError, LightningKokkosAMD,
                                                   QuantumFunctionError, Tape)
from pennylane_lightning_kokkos_amd.torch_adjoint import Op

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("c_dtype, tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
@pytest.mark.parametrize("par", [1, -2, 1.623, -0.051, 0])
def test_rx_ry_gradients(c_dtype, tol, par):
    """test_adjoint_jacobian.py:226-258: d<Z>/dx RX(x) = -sin x, d<X>/dy RY(y) = cos y."""
    dev = LightningKokkosAMD(1, c_dtype=c_dtype)
    jac = dev.adjoint_jacobian(Tape([Op("RX", [0], [par])], [("named", "PauliZ", [0])]))
    assert jac.shape == (1, 1) and abs(jac[0, 0] + np.sin(par)) < tol
    jac = dev.adjoint_jacobian(Tape([Op("RY", [0], [par])], [("named", "PauliX", [0])]))
    assert abs(jac[0, 0] - np.cos(par)) < tol


def test_rot_gradient_matches_oracle_and_fd():
    ops = [Op("RX", [0], [0.4]), Op("Rot", [1], [0.5, -0.3, 1.1]), Op("CNOT", [1, 2]),
           Op("CRY", [2, 0], [0.8]), Op("Rot", [2], [-0.2, 0.6, 0.9])]
    obs = [("named", "PauliZ", [0]), ("tensor", [("named", "PauliX", [1]), ("named", "PauliY", [2])])]
    dev = LightningKokkosAMD(3)
    tape = Tape(ops, obs)
    jac = dev.adjoint_jacobian(tape)
    flat = [p for o in ops for p in o.params]
    fd = np.zeros_like(jac)
    for k in range(len(flat)):
        vals = []
        for sgn in (1, -1):
            shifted, i = [], 0
            for o in ops:
                ps = []
                for p in o.params:
                    ps.append(p + (sgn * 1e-6 if i == k else 0.0))
                    i += 1
                shifted.append(Op(o.name, o.wires, ps))
            dev.reset()
            vals.append(dev.execute(Tape(shifted, obs)))
        fd[:, k] = (vals[0] - vals[1]) / 2e-6
    np.testing.assert_allclose(jac, fd, atol=1e-6)
    # the oracle on the serialized (Rot-expanded) circuit
    names, params, wires = [], [], []
    for o in ops:
        if o.name == "Rot":
            for g, a in zip(("RZ", "RY", "RZ"), o.params):
                names.append(g); params.append([a]); wires.append(o.wires)
        else:
            names.append(o.name); params.append(list(o.params)); wires.append(o.wires)
    ref = O.adjoint_jacobian(O.zero_state(3), [to_oracle(d) for d in obs],
                             O.Ops(names, params, wires, [False] * len(names), [[]] * len(names)),
                             list(range(len(flat))), True)
    np.testing.assert_allclose(jac, ref, atol=1e-12)


def test_stateprep_column_is_zero_and_others_shift():
    """lightning_kokkos.py:659-684: the StatePrep parameter gets a zero column,
    the operation parameters are shifted by one."""
    psi = np.array([1, 1j, 0.5, -0.5], dtype=np.complex128)
    psi /= np.linalg.norm(psi)
    ops = [Op("StatePrep", [0, 1], [psi]), Op("RX", [0], [0.3]), Op("RY", [1], [-0.7])]
    dev = LightningKokkosAMD(2)
    obs = [("named", "PauliZ", [0]), ("named", "PauliX", [1])]
    jac = dev.adjoint_jacobian(Tape(ops, obs, trainable_params=[0, 1, 2]))
    assert jac.shape == (2, 3) and np.all(jac[:, 0] == 0)
    ref = O.adjoint_jacobian(psi, [to_oracle(d) for d in obs],
                             O.Ops(["RX", "RY"], [[0.3], [-0.7]], [[0], [1]], [False] * 2, [[]] * 2),
                             [0, 1], True)
    np.testing.assert_allclose(jac[:, 1:], ref, atol=1e-12)


def test_vjp_is_dy_dot_jacobian():
    """test_vjp.py:41-71."""
    ops = [Op("RX", [0], [0.1]), Op("IsingXX", [0, 1], [0.7]), Op("Rot", [1], [0.2, 0.3, -0.4])]
    obs = [("named", "PauliZ", [0]), ("named", "PauliY", [1])]
    dev = LightningKokkosAMD(2)
    tape = Tape(ops, obs)
    jac = dev.adjoint_jacobian(tape)
    dy = np.array([0.6, -1.7])
    np.testing.assert_allclose(dev.vjp(obs, dy)(tape), dy @ jac, atol=1e-12)
    assert np.all(dev.vjp(obs, np.zeros(2))(tape) == 0)


def test_device_errors():
    dev = LightningKokkosAMD(2)
    with pytest.raises(QuantumFunctionError, match="Hermitian"):
        dev.adjoint_jacobian(Tape([Op("RX", [0], [0.1])], [("herm", np.eye(2), [0])]))
    with pytest.raises(QuantumFunctionError, match="CRot operation is not supported"):
        dev.adjoint_jacobian(Tape([Op("CRot", [0, 1], [0.1, 0.2, 0.3])], [("named", "PauliZ", [0])]))
    with pytest.raises(DeviceError, match="cannot be used after"):
        dev.apply([Op("RX", [0], [0.1]), Op("BasisState", [0, 1], [np.array([1, 0])])])
    with pytest.raises(RuntimeError, match="out-of-order indices"):
        dev.probability([1, 0])


def test_probability_and_samples():
    """test_probs.py:136-161 and test_sample.py:60-77."""
    dev = LightningKokkosAMD(2, shots=1000)
    dev.apply([Op("RX", [0], [0.5]), Op("RY", [1], [0.3])])
    np.testing.assert_allclose(dev.probability([0, 1]), [0.9178264236525453, 0.02096485729264079,
                                                         0.059841820910257436, 0.0013668981445561978],
                               atol=1e-12)
    np.testing.assert_allclose(dev.probability([0]), [0.938791280945186, 0.061208719054813635],
                               atol=1e-12)
    dev.reset()
    dev.apply([Op("RX", [0], [1.5708])])
    s = dev.generate_samples()
    assert s.shape == (1000, 2)
    z = 1 - 2 * s[:, 0]
    assert np.allclose(z ** 2, 1) and abs(z.mean()) < 0.15


def test_basis_state_and_subset_stateprep():
    dev = LightningKokkosAMD(3)
    dev.apply([Op("BasisState", [0, 2], [np.array([1, 1])])])
    st = dev.state
    assert st[0b101] == 1
    dev.apply([Op("StatePrep", [2], [np.array([0, 1])])])
    assert dev.state[0b001] == 1


@pytest.mark.parametrize("c_dtype, tol", [(np.complex128, 1e-12), (np.complex64, 2e-5)])
def test_batch_adjoint_jacobian_matches_per_tape(c_dtype, tol):
    """batch_adjoint_jacobian: eligible tapes (Rot / RY / CNOT / IsingXX, one
    Pauli-sum observable, trainable subsets) run in one batched launch, the
    others (a Toffoli, two observables, a state preparation, an explicit
    matrix, no trainable parameter) through adjoint_jacobian; every result
    equals the per-tape call."""
    rng = np.random.default_rng(9900)
    n = 6
    dev = LightningKokkosAMD(n, c_dtype=c_dtype)

    def rand_ops(k):
        ops = []
        for _ in range(k):
            g = ["Rot", "RY", "CNOT", "IsingXX", "RZ"][int(rng.integers(5))]
            if g == "Rot":
                ops.append(Op(g, [int(rng.integers(n))], [float(x) for x in rng.uniform(0, 6, 3)]))
            elif g in ("RY", "RZ"):
                ops.append(Op(g, [int(rng.integers(n))], [float(rng.uniform(0, 6))]))
            else:
                w = [int(x) for x in rng.choice(n, 2, replace=False)]
                ops.append(Op(g, w, [float(rng.uniform(0, 6))] if g == "IsingXX" else []))
        return ops

    ham = ("ham", [0.5, -1.2, 0.3], [("named", "PauliZ", [0]),
                                     ("tensor", [("named", "PauliX", [1]), ("named", "PauliY", [3])]),
                                     ("named", "PauliX", [5])])
    tapes = []
    for k in range(6):
        ops = rand_ops(10 + k)
        npar = sum(len(o.params) for o in ops)
        tp = None if k % 2 == 0 else sorted(int(x) for x in rng.choice(npar, npar // 2, replace=False))
        tapes.append(Tape(ops, [ham if k % 3 else ("named", "PauliZ", [k % n])], tp))
    tapes.append(Tape(rand_ops(8) + [Op("Toffoli", [0, 1, 2], [])], [ham]))
    tapes.append(Tape(rand_ops(8), [ham, ("named", "PauliZ", [2])]))
    tapes.append(Tape([Op("BasisState", [0, 1], [np.array([1, 0])])] + rand_ops(8), [ham]))
    q, _ = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))
    tapes.append(Tape(rand_ops(6) + [Op("QubitUnitary", [1, 4], [], False, q)], [ham]))
    tapes.append(Tape([Op("CNOT", [0, 1], [])], [ham]))
    got = dev.batch_adjoint_jacobian(tapes)
    for tape, g in zip(tapes, got):
        want = dev.adjoint_jacobian(tape)
        assert np.shape(g) == np.shape(want)
        np.testing.assert_allclose(g, want, atol=tol, rtol=0)
