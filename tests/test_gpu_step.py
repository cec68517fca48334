"""Single-op step kernels: operations the fused passes do not hold
(DoubleExcitation±, Toffoli, CSWAP, MultiRZ on 3+ wires, 3-4-wire matrices)
in the adjoint sweep and the forward application, against the oracle.

Each step touches only the local states its op moves (2 of 16 for
DoubleExcitation, 2 of 8 for Toffoli and CSWAP, padded subspaces for
matrices that are the identity on part of their states). Circuits mix those
ops with fused one- and two-wire gates so every path
around them runs too: register segments (8 qubits), register tiles (12 and
20 qubits complex128), LDS tiles (16 qubits), and HIP-graph replays. Nine
observables exercise two observable chunks of a step launch.
Tolerances: complex128 1e-10 relative + 1e-12 absolute per Jacobian entry,
1e-12 per amplitude; complex64 1e-5 (FP64 sums, FP32 amplitudes)."""
import numpy as np
import pytest

import oracle as O
from helpers import random_hamiltonian, random_unitary, to_capi, to_oracle

pytestmark = pytest.mark.gpu

WIDE = ["DoubleExcitation", "DoubleExcitationPlus", "DoubleExcitationMinus", "Toffoli", "CSWAP",
        "MultiRZ", "MultiRZ4", "MultiRZ6", "QubitUnitary3", "QubitUnitary4"]
NARROW = ["RX", "RY", "RZ", "CNOT", "IsingXX", "CRY", "Hadamard"]
NPAR = {"DoubleExcitation": 1, "DoubleExcitationPlus": 1, "DoubleExcitationMinus": 1,
        "MultiRZ": 1, "RX": 1, "RY": 1, "RZ": 1, "IsingXX": 1, "CRY": 1}
WIRES = {"DoubleExcitation": 4, "DoubleExcitationPlus": 4, "DoubleExcitationMinus": 4,
         "Toffoli": 3, "CSWAP": 3, "MultiRZ": 3, "RX": 1, "RY": 1, "RZ": 1, "CNOT": 2,
         "IsingXX": 2, "CRY": 2, "Hadamard": 1}


def _circuit(n, count, rng, wide_every=3):
    names, params, wires, inverses, mats = [], [], [], [], []
    for k in range(count):
        g = WIDE[rng.integers(len(WIDE))] if k % wide_every == 0 else NARROW[rng.integers(len(NARROW))]
        mat = []
        if g.startswith("QubitUnitary"):
            m = int(g[-1])
            g = "QubitUnitary"
            w = [int(x) for x in rng.choice(n, size=m, replace=False)]
            mat = random_unitary(2**m, rng).reshape(-1)
        elif g.startswith("MultiRZ"):
            m = int(g[7:]) if len(g) > 7 else 3
            g = "MultiRZ"
            w = [int(x) for x in rng.choice(n, size=min(m, n), replace=False)]
        else:
            w = [int(x) for x in rng.choice(n, size=WIRES[g], replace=False)]
        names.append(g)
        params.append([float(rng.uniform(0, 2 * np.pi))] if NPAR.get(g, 0) else [])
        wires.append(w)
        inverses.append(bool(rng.integers(2)))
        mats.append(mat)
    return names, params, wires, inverses, mats


def _obs(n, rng, count):
    out = [("named", "PauliZ", [int(rng.integers(n))]), random_hamiltonian(n, 6, rng)]
    while len(out) < count:
        out.append(("named", ("PauliX", "PauliY", "PauliZ")[len(out) % 3], [int(rng.integers(n))]))
    return out


def _tol(dtype):
    return dict(rtol=1e-10, atol=1e-12) if dtype == np.complex128 else dict(atol=1e-5)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,n_obs", [(8, 2), (12, 9), (16, 3), (20, 2)])
def test_adjoint_with_wide_ops(capi, monkeypatch, n, n_obs, dtype):
    monkeypatch.setenv("PLK_GRAPH", "0")
    rng = np.random.default_rng(9100 + n + n_obs)
    ops = _circuit(n, 45, rng)
    obs = _obs(n, rng, n_obs)
    npar = sum(1 for p in ops[1] if p)
    tp = list(range(npar)) if n <= 16 else list(range(npar - 10, npar))
    s = capi.State(n, dtype)
    capi.profiling_enable(True)
    try:
        got = capi.adjoint_jacobian(s, [to_capi(capi, o) for o in obs], capi.Ops(*ops), tp, True)
        steps = capi.profiling_get("adjoint_step")["launches"]
    finally:
        capi.profiling_enable(False)
    ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops), tp, True)
    np.testing.assert_allclose(got, ref, **_tol(dtype))
    assert steps > 0, "no single-op step kernel ran"


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [9, 13])
def test_forward_with_wide_ops(capi, n, dtype):
    rng = np.random.default_rng(9200 + n)
    ops = _circuit(n, 60, rng, wide_every=2)
    s = capi.State(n, dtype)
    s.apply_ops(capi.Ops(*ops))
    ref = O.apply_ops(O.zero_state(n), O.Ops(*ops))
    atol = 1e-12 if dtype == np.complex128 else 2e-5
    np.testing.assert_allclose(s.download(), ref, atol=atol)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_graph_replays_with_wide_ops(capi, monkeypatch, dtype):
    """Wide ops no longer keep a circuit out of graph capture: replays with
    new angles match the eager path and the oracle."""
    rng = np.random.default_rng(9300)
    n = 10
    base = _circuit(n, 40, rng)
    obs = _obs(n, rng, 3)
    cobs = [to_capi(capi, o) for o in obs]
    npar = sum(1 for p in base[1] if p)
    tp = list(range(npar))
    s = capi.State(n, dtype)
    for _ in range(4):
        params = [[float(rng.uniform(0, 2 * np.pi))] if p else [] for p in base[1]]
        ops = (base[0], params, base[2], base[3], base[4])
        monkeypatch.setenv("PLK_GRAPH", "1")
        got = capi.adjoint_jacobian(s, cobs, capi.Ops(*ops), tp, True)
        monkeypatch.setenv("PLK_GRAPH", "0")
        eager = capi.adjoint_jacobian(s, cobs, capi.Ops(*ops), tp, True)
        np.testing.assert_allclose(got, eager, rtol=1e-12, atol=1e-13)
        ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops), tp, True)
        np.testing.assert_allclose(got, ref, **_tol(dtype))


def test_wide_only_circuit_trainable_subset(capi, monkeypatch):
    """Only wide ops, a trainable subset and inverse flags: the step kernels
    carry the whole sweep (with early stop once the subset is consumed)."""
    monkeypatch.setenv("PLK_GRAPH", "0")
    rng = np.random.default_rng(9400)
    n = 11
    ops = _circuit(n, 30, rng, wide_every=1)
    obs = _obs(n, rng, 2)
    npar = sum(1 for p in ops[1] if p)
    tp = sorted(int(x) for x in rng.choice(npar, size=max(1, npar // 2), replace=False))
    s = capi.State(n)
    got = capi.adjoint_jacobian(s, [to_capi(capi, o) for o in obs], capi.Ops(*ops), tp, True)
    ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops), tp, True)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)


def _block_unitary(m, states, rng):
    """Identity on 2^m states except a random unitary on `states`."""
    U = np.eye(2**m, dtype=complex)
    B = random_unitary(len(states), rng)
    for a, r in enumerate(states):
        for b, c in enumerate(states):
            U[r, c] = B[a, b]
    return U


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_matrices_on_state_subspaces(capi, monkeypatch, dtype):
    """Matrices that move 2, 3, 4 or 5 of their local states (subspaces of
    2, 4, 4 and 8 after padding), forward and in the adjoint sweep."""
    monkeypatch.setenv("PLK_GRAPH", "0")
    rng = np.random.default_rng(9500)
    n = 13
    names, params, wires, inverses, mats = [], [], [], [], []
    blocks = [(3, [1, 6]), (4, [0, 5, 10]), (4, [3, 12, 7, 9]), (3, [0, 2, 4, 5, 7])]
    for k in range(24):
        m, states = blocks[k % len(blocks)]
        names.append("QubitUnitary")
        params.append([])
        wires.append([int(x) for x in rng.choice(n, size=m, replace=False)])
        inverses.append(bool(rng.integers(2)))
        mats.append(_block_unitary(m, states, rng).reshape(-1))
        g = ["RY", "RZ", "IsingXX"][k % 3]
        names.append(g)
        params.append([float(rng.uniform(0, 2 * np.pi))])
        wires.append([int(x) for x in rng.choice(n, size=2 if g == "IsingXX" else 1, replace=False)])
        inverses.append(False)
        mats.append([])
    ops = (names, params, wires, inverses, mats)
    s = capi.State(n, dtype)
    s.apply_ops(capi.Ops(*ops))
    ref = O.apply_ops(O.zero_state(n), O.Ops(*ops))
    np.testing.assert_allclose(s.download(), ref, atol=1e-12 if dtype == np.complex128 else 2e-5)
    obs = _obs(n, rng, 3)
    npar = sum(1 for p in params if p)
    fresh = capi.State(n, dtype)
    got = capi.adjoint_jacobian(fresh, [to_capi(capi, o) for o in obs], capi.Ops(*ops), list(range(npar)),
                                True)
    want = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops),
                              list(range(npar)), True)
    np.testing.assert_allclose(got, want, **_tol(dtype))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_excitation_workload(capi, dtype):
    """The benchmark's chemistry-style ansatz (20 qubits, 120 DoubleExcitation
    + 60 SingleExcitation, 32-word Hamiltonian + 4 PauliZ) against the oracle
    on the last 12 parameters and a few early ones."""
    from helpers import excitation_workload

    n, ops, obs = excitation_workload()
    npar = sum(1 for p in ops[1] if p)
    tp = [0, 1, 5] + list(range(npar - 12, npar))
    s = capi.State(n, dtype)
    got = capi.adjoint_jacobian(s, [to_capi(capi, o) for o in obs], capi.Ops(*ops), tp, True)
    ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops), tp, True)
    np.testing.assert_allclose(got, ref, **_tol(dtype))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_wide_matrices_between_side_by_side_passes(capi, monkeypatch, dtype):
    """5- and 6-wire matrices (the per-state path) between register-tile
    passes whose observable chunks run side by side, so lambda alternates
    between its two buffers around them; nine observables, against the
    oracle."""
    monkeypatch.setenv("PLK_GRAPH", "0")
    rng = np.random.default_rng(9600)
    n = 14
    names, params, wires, inverses, mats = [], [], [], [], []
    for k in range(30):
        if k % 10 == 5:
            m = 5 if k < 20 else 6
            names.append("QubitUnitary")
            params.append([])
            wires.append([int(x) for x in rng.choice(n, size=m, replace=False)])
            inverses.append(bool(rng.integers(2)))
            mats.append(random_unitary(2**m, rng).reshape(-1))
        else:
            g = ["RX", "RY", "CNOT", "IsingXX"][k % 4]
            names.append(g)
            params.append([float(rng.uniform(0, 2 * np.pi))] if g != "CNOT" else [])
            wires.append([int(x) for x in rng.choice(n, size=1 if g in ("RX", "RY") else 2,
                                                     replace=False)])
            inverses.append(False)
            mats.append([])
    ops = (names, params, wires, inverses, mats)
    obs = _obs(n, rng, 9)
    npar = sum(1 for p in params if p)
    got = capi.adjoint_jacobian(capi.State(n, dtype), [to_capi(capi, o) for o in obs], capi.Ops(*ops),
                                list(range(npar)), True)
    ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops),
                             list(range(npar)), True)
    np.testing.assert_allclose(got, ref, **_tol(dtype))
