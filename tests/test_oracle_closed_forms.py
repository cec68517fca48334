"""Pins of every gate and generator of the oracle by matrices written here
from their mathematical definitions, never from the oracle's own tables
(O.gate_matrix / O.generator_matrix are not called).

* gates: textbook matrices (wires[0] the most significant local bit); the
  controlled gates as |0><0| (x) I + |1><1| (x) U with U the single-qubit
  gate, control on wires[0];
* generators: G as the reference's generator kernels apply it in place and
  the scale s they return (GateFunctors.hpp:2172-3132,
  StateVectorKokkos.hpp:1275-1580), written as numpy matrices; the oracle's
  apply_generator must return s and G psi, and U(theta) must equal
  expm(i s theta G) for the gate of the first table.
"""
import numpy as np
import pytest
import scipy.linalg

import oracle as O
from helpers import GATES, GEN_GATES, random_state
from test_oracle_math import np_apply

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], complex)
Y = np.array([[0, -1j], [1j, 0]], complex)
Z = np.diag([1.0, -1.0]).astype(complex)
H = np.array([[1, 1], [1, -1]], complex) / np.sqrt(2)
P0 = np.diag([1.0, 0.0]).astype(complex)
P1 = np.diag([0.0, 1.0]).astype(complex)


def rx(t):
    return np.cos(t / 2) * I2 - 1j * np.sin(t / 2) * X


def ry(t):
    return np.cos(t / 2) * I2 - 1j * np.sin(t / 2) * Y


def rz(t):
    return np.cos(t / 2) * I2 - 1j * np.sin(t / 2) * Z


def phase(t):
    return np.diag([1.0, np.exp(1j * t)])


def rot(phi, theta, omega):
    return rz(omega) @ ry(theta) @ rz(phi)


def ctrl(U):
    """|0><0| (x) I + |1><1| (x) U, control = the more significant wire."""
    d = U.shape[0]
    return np.kron(P0, np.eye(d)) + np.kron(P1, U)


def pauli_exp(P, t):
    """exp(-i t/2 P) for a Pauli string P (P^2 = I)."""
    return np.cos(t / 2) * np.eye(P.shape[0]) - 1j * np.sin(t / 2) * P


SWAP = np.eye(4)[[0, 2, 1, 3]].astype(complex)


def givens(t, idx_a, idx_b, dim, other_phase):
    """Rotation by t/2 in the plane (a, b): a -> cos a + sin b, others times other_phase."""
    M = np.eye(dim, dtype=complex) * other_phase
    c, s = np.cos(t / 2), np.sin(t / 2)
    M[idx_a, idx_a] = c
    M[idx_b, idx_b] = c
    M[idx_b, idx_a] = s
    M[idx_a, idx_b] = -s
    return M


def gate(name, params, nw):
    t = params[0] if params else None
    table = {
        "Identity": lambda: I2,
        "PauliX": lambda: X, "PauliY": lambda: Y, "PauliZ": lambda: Z, "Hadamard": lambda: H,
        "S": lambda: np.diag([1, 1j]), "T": lambda: np.diag([1, np.exp(1j * np.pi / 4)]),
        "PhaseShift": lambda: phase(t), "RX": lambda: rx(t), "RY": lambda: ry(t), "RZ": lambda: rz(t),
        "Rot": lambda: rot(*params),
        "CNOT": lambda: ctrl(X), "CY": lambda: ctrl(Y), "CZ": lambda: ctrl(Z), "SWAP": lambda: SWAP,
        "ControlledPhaseShift": lambda: ctrl(phase(t)), "CRX": lambda: ctrl(rx(t)),
        "CRY": lambda: ctrl(ry(t)), "CRZ": lambda: ctrl(rz(t)), "CRot": lambda: ctrl(rot(*params)),
        "IsingXX": lambda: pauli_exp(np.kron(X, X), t), "IsingYY": lambda: pauli_exp(np.kron(Y, Y), t),
        "IsingZZ": lambda: pauli_exp(np.kron(Z, Z), t),
        # exp(i t/4 (XX + YY)): rotation by t/2 mixing |01> and |10> with i sin.
        "IsingXY": lambda: scipy.linalg.expm(0.25j * t * (np.kron(X, X) + np.kron(Y, Y))),
        "SingleExcitation": lambda: givens(t, 1, 2, 4, 1.0),
        "SingleExcitationMinus": lambda: _se_pm(t, -1),
        "SingleExcitationPlus": lambda: _se_pm(t, +1),
        "CSWAP": lambda: np.kron(P0, np.eye(4)) + np.kron(P1, SWAP),
        "Toffoli": lambda: ctrl(ctrl(X)),
        "DoubleExcitation": lambda: givens(t, 3, 12, 16, 1.0),
        "DoubleExcitationMinus": lambda: givens(t, 3, 12, 16, np.exp(-0.5j * t)),
        "DoubleExcitationPlus": lambda: givens(t, 3, 12, 16, np.exp(0.5j * t)),
        "MultiRZ": lambda: np.diag([np.exp(-0.5j * t * (1 - 2 * (bin(k).count("1") % 2))) for k in range(2**nw)]),
    }
    return np.asarray(table[name](), dtype=complex)


def _se_pm(t, sign):
    M = givens(t, 1, 2, 4, 1.0)
    M[0, 0] = M[3, 3] = np.exp(sign * 0.5j * t)
    return M


def generator(name, nw):
    """(G as applied in place by the reference's generator kernel, scale s)."""
    def emb(Gsub, idx, dim, rest=0.0):
        G = np.eye(dim, dtype=complex) * rest
        for a, ra in enumerate(idx):
            for b, rb in enumerate(idx):
                G[ra, rb] = Gsub[a, b]
        return G

    parity = np.diag([1.0 - 2 * (bin(k).count("1") % 2) for k in range(2**nw)]).astype(complex)
    table = {
        "RX": (X, -0.5), "RY": (Y, -0.5), "RZ": (Z, -0.5),
        "PhaseShift": (P1, 1.0),
        "ControlledPhaseShift": (np.kron(P1, P1), 1.0),
        "CRX": (np.kron(P1, X), -0.5), "CRY": (np.kron(P1, Y), -0.5), "CRZ": (np.kron(P1, Z), -0.5),
        "IsingXX": (np.kron(X, X), -0.5), "IsingYY": (np.kron(Y, Y), -0.5), "IsingZZ": (np.kron(Z, Z), -0.5),
        # Swap of |01> and |10>, |00> and |11> zeroed.
        "IsingXY": (emb(X, [1, 2], 4), 0.5),
        # |01> <- -i |10>, |10> <- i |01> (Y on the pair); |00>, |11> zeroed / kept / negated.
        "SingleExcitation": (emb(Y, [1, 2], 4), -0.5),
        "SingleExcitationMinus": (emb(Y, [1, 2], 4, rest=1.0), -0.5),
        "SingleExcitationPlus": (emb(Y, [1, 2], 4, rest=-1.0), -0.5),
        # |0011> <- -i |1100>, |1100> <- i |0011>; every other amplitude zeroed / kept.
        "DoubleExcitation": (emb(Y, [3, 12], 16), -0.5),
        "DoubleExcitationMinus": (emb(Y, [3, 12], 16, rest=1.0), -0.5),
        # |0011> <- i |1100>, |1100> <- -i |0011>; every other amplitude kept.
        "DoubleExcitationPlus": (emb(-Y, [3, 12], 16, rest=1.0), 0.5),
        "MultiRZ": (parity, -0.5),
    }
    G, s = table[name]
    return np.asarray(G, dtype=complex), s


def _params(name, rng):
    return [float(x) for x in rng.uniform(-3, 3, size=GATES[name][1])]


@pytest.mark.parametrize("name", list(GATES))
def test_gate_equals_closed_form(name):
    rng = np.random.default_rng(sum(map(ord, name)))
    n = 5
    nw = GATES[name][0] or 3
    for trial in range(3):
        wires = [int(x) for x in rng.choice(n, size=nw, replace=False)]
        params = _params(name, rng)
        U = gate(name, params, nw)
        psi = random_state(n, rng)
        for inv, M in ((False, U), (True, U.conj().T)):
            got = O.apply(psi.copy(), name, wires, inv, params)
            np.testing.assert_allclose(got, np_apply(psi, M, wires), atol=1e-12,
                                       err_msg=f"{name} wires={wires} inverse={inv}")


@pytest.mark.parametrize("name", GEN_GATES)
def test_generator_equals_closed_form(name):
    rng = np.random.default_rng(100 + sum(map(ord, name)))
    n = 5
    nw = GATES[name][0] or 3
    G, s = generator(name, nw)
    for _ in range(3):
        wires = [int(x) for x in rng.choice(n, size=nw, replace=False)]
        psi = random_state(n, rng)
        out = psi.copy()
        got_s = O.apply_generator(out, name, wires)
        assert got_s == s
        np.testing.assert_allclose(out, np_apply(psi, G, wires), atol=1e-13, err_msg=f"{name} wires={wires}")


@pytest.mark.parametrize("name", GEN_GATES)
def test_closed_form_gate_is_exp_of_closed_form_generator(name):
    """Ties the two tables above together: U(theta) = expm(i s theta G)."""
    nw = GATES[name][0] or 3
    G, s = generator(name, nw)
    for theta in (0.0, 0.41, -2.2, 3.0):
        np.testing.assert_allclose(gate(name, [theta], nw), scipy.linalg.expm(1j * s * theta * G), atol=1e-12)


def test_every_gate_and_generator_has_an_independent_pin():
    assert set(GATES) <= {n for n in GATES if gate(n, [0.3] * GATES[n][1], GATES[n][0] or 3) is not None}
    for name in GEN_GATES:
        generator(name, GATES[name][0] or 3)
