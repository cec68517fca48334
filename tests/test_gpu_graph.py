"""Adjoint sweeps replayed from HIP graphs (circuits of up to 18 qubits with
Pauli-word observables) against the eager launches and the oracle.

A graph is captured once per circuit structure and replayed with freshly
uploaded plan tables, so every test replays it several times with new
angles, with other eager work and synchronisations in between, and checks
each replay: against the eager path (PLK_GRAPH=0) to 1e-12 and against the
oracle to the usual 1e-10 relative + 1e-12 absolute (complex128), 1e-5
(complex64)."""
import numpy as np
import pytest

import oracle as O
from helpers import c1_workload, random_hamiltonian, random_ops, random_state, to_capi, to_oracle

pytestmark = pytest.mark.gpu

NAMES = ["RX", "RY", "RZ", "PhaseShift", "CNOT", "CZ", "IsingXX", "IsingZZ", "CRY", "Hadamard", "SWAP"]


def _tol(dtype):
    return dict(rtol=1e-10, atol=1e-12) if dtype == np.complex128 else dict(atol=1e-5)


def _reparam(ops, rng, value=None):
    params = [[float(value if value is not None else rng.uniform(0, 2 * np.pi))] if p else [] for p in ops[1]]
    return (ops[0], params, ops[2], ops[3], ops[4])


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [6, 10, 14])
def test_graph_replays_match_eager_and_oracle(capi, monkeypatch, n, dtype):
    rng = np.random.default_rng(7700 + n)
    base = random_ops(n, 60, rng, names=NAMES)
    obs = [("named", "PauliZ", [0]), random_hamiltonian(n, 5, rng), ("named", "PauliX", [n - 1])]
    cobs = [to_capi(capi, o) for o in obs]
    npar = sum(1 for p in base[1] if p)
    tp = list(range(0, npar, 2))
    s = capi.State(n, dtype)
    other = capi.State(n, dtype)
    for k in range(4):
        ops = _reparam(base, rng)
        cops = capi.Ops(*ops)
        monkeypatch.setenv("PLK_GRAPH", "1")
        got = capi.adjoint_jacobian(s, cobs, cops, tp, True)
        # Eager work between replays.
        other.apply_ops(cops)
        other.expval(cobs[1])
        monkeypatch.setenv("PLK_GRAPH", "0")
        eager = capi.adjoint_jacobian(s, cobs, cops, tp, True)
        np.testing.assert_allclose(got, eager, rtol=1e-12, atol=1e-13)
        ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops), tp, True)
        np.testing.assert_allclose(got, ref, **_tol(dtype))


def test_graph_layout_change_recaptures(capi, monkeypatch):
    """All angles 0 make RX / RY diagonal and change the packed tables'
    layout; the next call with generic angles must not replay that graph."""
    monkeypatch.setenv("PLK_GRAPH", "1")
    rng = np.random.default_rng(7800)
    n = 9
    base = random_ops(n, 40, rng, names=["RX", "RY", "RZ", "CNOT"])
    obs = [("named", "PauliZ", [w]) for w in range(3)]
    cobs = [to_capi(capi, o) for o in obs]
    npar = sum(1 for p in base[1] if p)
    s = capi.State(n)
    for value in (0.0, None, 0.0, None):
        ops = _reparam(base, rng, value)
        got = capi.adjoint_jacobian(s, cobs, capi.Ops(*ops), list(range(npar)), True)
        ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops),
                                 list(range(npar)), True)
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)


def test_graph_input_state_changes(capi, monkeypatch):
    """apply_operations = False: the replay copies the state's current
    contents (same buffer, new amplitudes)."""
    monkeypatch.setenv("PLK_GRAPH", "1")
    rng = np.random.default_rng(7900)
    n = 8
    ops = random_ops(n, 30, rng, names=NAMES)
    obs = [random_hamiltonian(n, 4, rng)]
    cobs = [to_capi(capi, o) for o in obs]
    npar = sum(1 for p in ops[1] if p)
    s = capi.State(n)
    for _ in range(3):
        psi = random_state(n, rng)
        s.upload(psi)
        got = capi.adjoint_jacobian(s, cobs, capi.Ops(*ops), list(range(npar)), False)
        ref = O.adjoint_jacobian(psi, [to_oracle(o) for o in obs], O.Ops(*ops), list(range(npar)), False)
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)


def test_graph_many_structures(capi, monkeypatch):
    """Many structures, cycled three times: the first call of each runs
    eagerly, the second captures, the third replays."""
    monkeypatch.setenv("PLK_GRAPH", "1")
    rng = np.random.default_rng(8000)
    n = 7
    circuits = [random_ops(n, 12 + k, rng, names=NAMES) for k in range(11)]
    obs = [("named", "PauliY", [1])]
    cobs = [to_capi(capi, o) for o in obs]
    s = capi.State(n)
    for _ in range(3):
        for ops in circuits:
            npar = sum(1 for p in ops[1] if p)
            if npar == 0:
                continue
            got = capi.adjoint_jacobian(s, cobs, capi.Ops(*ops), list(range(npar)), True)
            ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops),
                                     list(range(npar)), True)
            np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)


def test_graph_c1_workload(capi, monkeypatch):
    monkeypatch.setenv("PLK_GRAPH", "1")
    n, ops, obs = c1_workload()
    cobs = [to_capi(capi, o) for o in obs]
    npar = sum(1 for p in ops[1] if p)
    s = capi.State(n)
    rng = np.random.default_rng(8100)
    for _ in range(3):
        o2 = _reparam(ops, rng)
        got = capi.adjoint_jacobian(s, cobs, capi.Ops(*o2), list(range(npar)), True)
        ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*o2),
                                 list(range(npar)), True)
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)
