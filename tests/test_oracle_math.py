"""Pins of the oracle that do not rely on the oracle's own formulas.

* every named gate applied by the oracle equals a numpy tensordot of its
  matrix on the state (brute force; catches wire-order and bit-order slips);
* every parametric gate equals expm(i * s * theta * G) (scipy), with G and s
  from the oracle's generator table; the inverse flag gives U^dagger;
* the adjoint Jacobian equals central finite differences (h = 1e-6, FP64)
  and, for Pauli rotations, the exact parameter-shift rule;
* expectation values equal dense numpy <psi|O|psi>.
"""
import numpy as np
import pytest
import scipy.linalg

import oracle as O
from helpers import (GATES, GEN_GATES, random_hamiltonian, random_hermitian, random_ops, random_state,
                     random_unitary, to_oracle)

PAULI = {"PauliX": np.array([[0, 1], [1, 0]], complex), "PauliY": np.array([[0, -1j], [1j, 0]], complex),
         "PauliZ": np.diag([1.0, -1.0]).astype(complex), "Identity": np.eye(2, dtype=complex),
         "Hadamard": np.array([[1, 1], [1, -1]], complex) / np.sqrt(2)}


def np_apply(psi, U, wires):
    """Applies U (2^m x 2^m, wires[0] most significant) with tensordot; wire w is axis w."""
    n = int(np.log2(psi.size))
    m = len(wires)
    t = psi.reshape([2] * n)
    Ut = U.reshape([2] * (2 * m))
    out = np.tensordot(Ut, t, axes=(list(range(m, 2 * m)), list(wires)))
    out = np.moveaxis(out, list(range(m)), list(wires))
    return out.reshape(-1)


def dense_obs(ob, n):
    """Dense 2^n x 2^n matrix of an observable description."""
    kind = ob[0]
    if kind == "named":
        return full_op(PAULI[ob[1]], ob[2], n)
    if kind == "herm":
        return full_op(np.asarray(ob[1]), ob[2], n)
    if kind == "tensor":
        M = np.eye(2**n, dtype=complex)
        for f in ob[1]:
            M = dense_obs(f, n) @ M
        return M
    if kind == "ham":
        return sum(c * dense_obs(t, n) for c, t in zip(ob[1], ob[2]))
    raise ValueError(kind)


def full_op(U, wires, n):
    cols = [np_apply(e, U, wires) for e in np.eye(2**n, dtype=complex)]
    return np.array(cols).T


@pytest.mark.parametrize("name", [g for g in GATES if g not in ("Identity",)])
def test_named_gate_equals_tensordot(name):
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    n = 5
    nw, npar = GATES[name]
    nw = nw or 3
    for inv in (False, True):
        wires = [int(x) for x in rng.choice(n, size=nw, replace=False)]
        params = list(rng.uniform(-3, 3, size=npar))
        U = O.gate_matrix(name, nw, params, inv)
        assert np.allclose(U @ U.conj().T, np.eye(2**nw), atol=1e-12), "not unitary"
        psi = random_state(n, rng)
        got = O.apply(psi.copy(), name, wires, inv, params)
        np.testing.assert_allclose(got, np_apply(psi, U, wires), atol=1e-12)


@pytest.mark.parametrize("name", [g for g in GATES if GATES[g][1] >= 1])
def test_inverse_flag_is_dagger(name):
    nw = GATES[name][0] or 3
    params = list(np.random.default_rng(3).uniform(-3, 3, size=GATES[name][1]))
    U = O.gate_matrix(name, nw, params, False)
    Ui = O.gate_matrix(name, nw, params, True)
    np.testing.assert_allclose(Ui, U.conj().T, atol=1e-12)


@pytest.mark.parametrize("name", GEN_GATES)
def test_gate_is_exp_of_generator(name):
    nw = GATES[name][0] or 3
    G, s = O.generator_matrix(name, nw)
    for theta in (0.0, 0.37, -1.9, 2.8):
        U = O.gate_matrix(name, nw, [theta], False)
        np.testing.assert_allclose(U, scipy.linalg.expm(1j * s * theta * G), atol=1e-12)


@pytest.mark.parametrize("name", GEN_GATES)
def test_generator_central_difference(name):
    """Test_StateVectorKokkos_Generator.cpp pattern, tightened to FP64."""
    rng = np.random.default_rng(11)
    n = 4
    nw = GATES[name][0] or 3
    wires = [int(x) for x in rng.choice(n, size=nw, replace=False)]
    psi = random_state(n, rng)
    eps = 1e-6
    plus = O.apply(psi.copy(), name, wires, False, [eps])
    minus = O.apply(psi.copy(), name, wires, False, [-eps])
    gpsi = psi.copy()
    s = O.apply_generator(gpsi, name, wires)
    np.testing.assert_allclose((plus - minus) / (2 * eps), 1j * s * gpsi, atol=1e-8)


def _expvals(psi, obs):
    return np.array([O.expval(psi, to_oracle(o)) for o in obs])


@pytest.mark.parametrize("seed", range(4))
def test_adjoint_equals_central_fd(seed):
    rng = np.random.default_rng(50 + seed)
    n = 4
    names, params, wires, inv, mats = random_ops(n, 14, rng, with_matrix=(seed % 2 == 0))
    obs = [("named", "PauliZ", [0]), random_hamiltonian(n, 4, rng),
           ("herm", random_hermitian(4, rng), [1, 3])]
    psi0 = random_state(n, rng)
    par_idx = [i for i, p in enumerate(params) if p]
    jac = O.adjoint_jacobian(psi0, [to_oracle(o) for o in obs], O.Ops(names, params, wires, inv, mats),
                             list(range(len(par_idx))), True)
    h = 1e-6
    for col, i in enumerate(par_idx):
        vals = []
        for d in (h, -h):
            pp = [list(p) for p in params]
            pp[i] = [pp[i][0] + d]
            psi = O.apply_ops(psi0.copy(), O.Ops(names, pp, wires, inv, mats))
            vals.append(_expvals(psi, obs))
        fd = (vals[0] - vals[1]) / (2 * h)
        np.testing.assert_allclose(jac[:, col], fd, atol=2e-8, err_msg=f"param {i} ({names[i]})")


def test_adjoint_equals_parameter_shift():
    rng = np.random.default_rng(77)
    n = 5
    names, params, wires, inv, mats = random_ops(n, 20, rng, names=["RX", "RY", "RZ", "CNOT", "CZ", "Hadamard"],
                                                 allow_inverse=False)
    obs = [random_hamiltonian(n, 5, rng)]
    psi0 = O.zero_state(n)
    par_idx = [i for i, p in enumerate(params) if p]
    jac = O.adjoint_jacobian(psi0, [to_oracle(o) for o in obs], O.Ops(names, params, wires, inv, mats),
                             list(range(len(par_idx))), True)
    for col, i in enumerate(par_idx):
        vals = []
        for d in (np.pi / 2, -np.pi / 2):
            pp = [list(p) for p in params]
            pp[i] = [pp[i][0] + d]
            vals.append(_expvals(O.apply_ops(psi0.copy(), O.Ops(names, pp, wires, inv, mats)), obs))
        np.testing.assert_allclose(jac[:, col], (vals[0] - vals[1]) / 2, atol=1e-12)


def test_adjoint_tp_subset_and_inverse_sign():
    """Trainable subsets and the (inverse ? -1 : 1) sign of AdjointDiffKokkos.hpp:459-463."""
    rng = np.random.default_rng(8)
    n = 3
    names = ["RX", "IsingXX", "CRY", "RZ", "SingleExcitationPlus", "PhaseShift"]
    wires = [[0], [1, 2], [2, 0], [1], [0, 1], [2]]
    params = [[float(x)] for x in rng.uniform(0, 6, size=6)]
    obs = [("named", "PauliY", [1]), ("tensor", [("named", "PauliX", [0]), ("named", "PauliZ", [2])])]
    psi = random_state(n, rng)
    inv = [True, False, True, False, True, True]
    full = O.adjoint_jacobian(psi, [to_oracle(o) for o in obs], O.Ops(names, params, wires, inv), list(range(6)))
    sub = O.adjoint_jacobian(psi, [to_oracle(o) for o in obs], O.Ops(names, params, wires, inv), [1, 4, 5])
    np.testing.assert_allclose(sub, full[:, [1, 4, 5]], atol=1e-14)
    # Inverting a gate and negating its angle is the same operation: the
    # derivative w.r.t. the stored parameter flips sign.
    params_neg = [[-p[0]] for p in params]
    flipped = O.adjoint_jacobian(psi, [to_oracle(o) for o in obs],
                                 O.Ops(names, params_neg, wires, [not b for b in inv]), list(range(6)))
    np.testing.assert_allclose(flipped, -full, atol=1e-12)


def test_expval_equals_dense():
    rng = np.random.default_rng(9)
    n = 5
    psi = random_state(n, rng)
    cases = [("named", "Hadamard", [3]), ("tensor", [("named", "PauliY", [0]), ("named", "Hadamard", [4])]),
             random_hamiltonian(n, 7, rng), ("herm", random_hermitian(8, rng), [4, 0, 2])]
    for ob in cases:
        ref = np.vdot(psi, dense_obs(ob, n) @ psi).real
        assert O.expval(psi, to_oracle(ob)) == pytest.approx(ref, abs=1e-12)


def test_matrix_op_equals_tensordot():
    rng = np.random.default_rng(10)
    for k in (1, 2, 3, 4):
        U = random_unitary(2**k, rng)
        wires = [int(x) for x in rng.choice(6, size=k, replace=False)]
        psi = random_state(6, rng)
        np.testing.assert_allclose(O.apply_matrix(psi.copy(), U, wires), np_apply(psi, U, wires), atol=1e-12)
        np.testing.assert_allclose(O.apply_matrix(psi.copy(), U, wires, True),
                                   np_apply(psi, U.conj().T, wires), atol=1e-12)
