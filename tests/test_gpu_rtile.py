"""Register-tile passes against the oracle.

PLK_RTILE=1 forces register-tile passes from the smallest state they can
tile (9 qubits complex128, 10 complex64), so small random circuits exercise
every instruction shape: operations on register bits, diagonal operations
on lane bits, controlled operations with the control on a lane, register /
lane remaps, trainable operations of every generator kind, observable
chunks and both precisions. Tolerances as in the other parity tests:
complex128 1e-10 relative + 1e-12 absolute per Jacobian entry, 1e-12 per
amplitude; complex64 1e-5 absolute.
"""
import numpy as np
import pytest

import oracle as O
from helpers import GATES, random_hamiltonian, random_ops, random_state, to_capi, to_oracle

pytestmark = pytest.mark.gpu

ONE_TWO_WIRE = [g for g in GATES if g != "Identity" and GATES[g][0] in (1, 2) and GATES[g][1] <= 1]


@pytest.fixture
def rtile(monkeypatch):
    monkeypatch.setenv("PLK_RTILE", "1")


def _assert_jac(ref, got, dtype):
    if dtype == np.complex128:
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)
    else:
        np.testing.assert_allclose(got, ref, atol=1e-5)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("seed", range(4))
def test_rtile_adjoint_random(capi, rtile, seed, dtype):
    rng = np.random.default_rng(900 + seed)
    n = [10, 11, 12, 13][seed]
    ops = random_ops(n, 90, rng, names=ONE_TWO_WIRE, with_matrix=(seed % 2 == 1))
    npar = sum(1 for p in ops[1] if p)
    tp = sorted(int(x) for x in rng.choice(npar, size=max(1, npar - 5 * seed), replace=False))
    obs = [("named", "PauliZ", [j % n]) for j in range(1 + seed)] + [random_hamiltonian(n, 5, rng)]
    psi = random_state(n, rng)
    ref = O.adjoint_jacobian(psi, [to_oracle(o) for o in obs], O.Ops(*ops), tp, True)
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    got = capi.adjoint_jacobian(s, [to_capi(capi, o) for o in obs], capi.Ops(*ops), tp, True)
    _assert_jac(ref, got, dtype)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [10, 12, 14])
def test_rtile_adjoint_layered(capi, rtile, n, dtype):
    """Rot layers (RZ, RY, RZ) and CNOT ranges: the benchmark circuit's structure."""
    rng = np.random.default_rng(950 + n)
    names, params, wires = [], [], []
    for layer in range(3):
        for w in range(n):
            for g in ("RZ", "RY", "RZ"):
                names.append(g)
                params.append([float(rng.uniform(0, 2 * np.pi))])
                wires.append([w])
        r = (layer % (n - 1)) + 1
        for w in range(n):
            names.append("CNOT")
            params.append([])
            wires.append([w, (w + r) % n])
    ops = (names, params, wires, [False] * len(names), [[] for _ in names])
    obs = [("named", "PauliZ", [w]) for w in range(4)] + [random_hamiltonian(n, 8, rng)]
    npar = sum(1 for p in params if p)
    ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops), list(range(npar)), True)
    s = capi.State(n, dtype)
    got = capi.adjoint_jacobian(s, [to_capi(capi, o) for o in obs], capi.Ops(*ops), list(range(npar)), True)
    _assert_jac(ref, got, dtype)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [10, 13])
def test_rtile_forward(capi, rtile, n, dtype):
    rng = np.random.default_rng(980 + n)
    ops = random_ops(n, 160, rng, with_matrix=True)
    psi = random_state(n, rng)
    ref = O.apply_ops(psi.copy(), O.Ops(*ops))
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    s.apply_ops(capi.Ops(*ops))
    tol = 1e-12 if dtype == np.complex128 else 5e-5
    np.testing.assert_allclose(s.download(), ref, atol=tol)


def test_rtile_matches_lds_tiles(capi, monkeypatch):
    """At 18 qubits the register-tile and LDS-tile passes give the same Jacobian."""
    rng = np.random.default_rng(990)
    n = 18
    ops = random_ops(n, 120, rng, names=ONE_TWO_WIRE)
    obs = [("named", "PauliZ", [3]), random_hamiltonian(n, 4, rng), ("named", "PauliX", [n - 1])]
    npar = sum(1 for p in ops[1] if p)
    psi = random_state(n, rng)
    s = capi.State(n)
    cobs = [to_capi(capi, o) for o in obs]
    cops = capi.Ops(*ops)
    monkeypatch.setenv("PLK_RTILE", "1")
    s.upload(psi)
    a = capi.adjoint_jacobian(s, cobs, cops, list(range(npar)), True)
    monkeypatch.setenv("PLK_RTILE", "0")
    s.upload(psi)
    b = capi.adjoint_jacobian(s, cobs, cops, list(range(npar)), True)
    np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,nj", [(12, "3"), (12, "1"), (18, "")])
@pytest.mark.parametrize("n_obs", [2, 4, 5, 7, 8, 10, 25])
def test_side_by_side_chunks(capi, monkeypatch, n, nj, n_obs, dtype):
    """A pass runs its observable chunks side by side in one launch: chunks
    of nj and nj - 1 states (5 observables in chunks of 3 -> 3 + 2), launches
    split at 24 states (25 observables), lambda alternating between two
    buffers; against the oracle."""
    if nj:
        monkeypatch.setenv("PLK_RTILE_NJ", nj)
    monkeypatch.setenv("PLK_GRAPH", "0")
    rng = np.random.default_rng(9700 + 31 * n + n_obs)
    ops = random_ops(n, 50, rng, names=ONE_TWO_WIRE)
    obs = [("named", ("PauliX", "PauliY", "PauliZ")[k % 3], [int(rng.integers(n))]) for k in range(n_obs - 1)]
    obs.append(random_hamiltonian(n, 4, rng))
    npar = sum(1 for p in ops[1] if p)
    tp = list(range(npar)) if n <= 12 else list(range(npar - 12, npar))
    s = capi.State(n, dtype)
    got = capi.adjoint_jacobian(s, [to_capi(capi, o) for o in obs], capi.Ops(*ops), tp, True)
    ref = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops), tp, True)
    _assert_jac(ref, got, dtype)


SELECTOR_GATES = ["CNOT", "CZ", "CY", "CRX", "CRY", "CRZ", "ControlledPhaseShift", "IsingZZ", "RZ",
                  "PhaseShift", "MultiRZ", "SWAP"]


def _selector_circuit(n, count, rng):
    """Gates whose control or diagonal wires sit on the highest wires (low
    index bits, usually inside the tile) or the lowest wires (high index
    bits, usually outside it), mixed with RY / RX on the other end."""
    names, params, wires = [], [], []
    for k in range(count):
        g = SELECTOR_GATES[int(rng.integers(len(SELECTOR_GATES)))] if k % 2 else ["RY", "RX"][k % 4 // 2]
        lo = [int(x) for x in rng.choice(4, size=2, replace=False)]          # wires 0-3: high bits
        hi = [n - 1 - int(x) for x in rng.choice(4, size=2, replace=False)]  # wires n-4..n-1: low bits
        if g in ("RY", "RX"):
            w = [hi[0] if k % 3 else lo[0]]
        elif g in ("RZ", "PhaseShift"):
            w = [lo[0]]
        elif g == "SWAP":
            w = [lo[0], hi[0]]
        else:
            w = [lo[0], hi[0]] if k % 3 else [hi[0], lo[0]]
        names.append(g)
        params.append([float(rng.uniform(0, 2 * np.pi))] if GATES[g][1] == 1 else [])
        wires.append(w)
    return names, params, wires, [bool(rng.integers(2)) for _ in names], [[] for _ in names]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [16, 20])
def test_selectors_outside_the_tile(capi, monkeypatch, n, dtype):
    """Controls and diagonal factors on wires outside the tile select their
    block from the tile's fixed index bits: forward application and the
    adjoint against the oracle."""
    monkeypatch.setenv("PLK_GRAPH", "0")
    rng = np.random.default_rng(9800 + n)
    ops = _selector_circuit(n, 80, rng)
    s = capi.State(n, dtype)
    s.apply_ops(capi.Ops(*ops))
    ref = O.apply_ops(O.zero_state(n), O.Ops(*ops))
    np.testing.assert_allclose(s.download(), ref, atol=1e-12 if dtype == np.complex128 else 2e-5)
    obs = [("named", "PauliZ", [0]), ("named", "PauliX", [n - 1]), random_hamiltonian(n, 5, rng),
           ("tensor", [("named", "PauliY", [1]), ("named", "PauliZ", [n - 2])])]
    npar = sum(1 for p in ops[1] if p)
    tp = list(range(npar)) if n <= 16 else list(range(npar - 16, npar))
    got = capi.adjoint_jacobian(capi.State(n, dtype), [to_capi(capi, o) for o in obs], capi.Ops(*ops),
                                tp, True)
    want = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ops), tp, True)
    _assert_jac(want, got, dtype)
