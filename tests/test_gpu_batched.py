"""Batched adjoint Jacobians (plk_adjoint_batched) against the oracle.

Every circuit of a batch starts from |0...0>; its expectation value and
Jacobian row must equal the oracle's adjoint_jacobian with apply_operations
(the reference's AdjointDiffKokkos.hpp:404-478 restated on the CPU) and the
per-circuit engine entry point. complex128: 1e-10 relative; complex64 state
with FP64 accumulation: 1e-5 absolute. Small tile widths (PLK_BATCH_TILE_BITS)
force circuits out of LDS residency onto the streamed global path."""
import numpy as np
import pytest

import oracle as O
from helpers import GATES, c4_workload, random_hamiltonian, random_ops, to_capi, to_oracle

pytestmark = pytest.mark.gpu

TWO_WIRE = ["RX", "RY", "RZ", "PhaseShift", "Hadamard", "PauliX", "PauliY", "PauliZ", "S", "T",
            "CNOT", "CY", "CZ", "SWAP", "ControlledPhaseShift", "CRX", "CRY", "CRZ", "IsingXX",
            "IsingXY", "IsingYY", "IsingZZ", "SingleExcitation", "SingleExcitationMinus",
            "SingleExcitationPlus", "MultiRZ"]


def _oracle(n, ops, ob, tp):
    psi = O.zero_state(n)
    o = O.Ops(*ops)
    jac = O.adjoint_jacobian(psi, [to_oracle(ob)], o, tp, True)[0]
    e = O.expval(O.apply_ops(psi.copy(), o), to_oracle(ob))
    return e, jac


def _batched(capi, circuits, tps, dtype):
    ops = [capi.Ops(*c[1]) for c in circuits]
    obs = [to_capi(capi, c[2]) for c in circuits]
    return capi.adjoint_batched([c[0] for c in circuits], ops, obs, tps, dtype=dtype)


def _check(capi, circuits, tps, dtype):
    ev, rows = _batched(capi, circuits, tps, dtype)
    tol = dict(rtol=1e-10, atol=1e-12) if dtype == np.complex128 else dict(atol=1e-5)
    for i, (n, ops, ob) in enumerate(circuits):
        tp = tps[i] if tps is not None else list(range(sum(1 for p in ops[1] if p)))
        e, jac = _oracle(n, ops, ob, tp)
        np.testing.assert_allclose(ev[i], e, err_msg=f"circuit {i} n={n} expval", **tol)
        np.testing.assert_allclose(rows[i], jac, err_msg=f"circuit {i} n={n} jac", **tol)
    return ev, rows


def _random_batch(rng, count, nmin, nmax, nops=30):
    circuits, tps = [], []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        names = [g for g in TWO_WIRE if GATES[g][0] <= n]
        ops = random_ops(n, nops, rng, names=names)
        # MultiRZ on at most two wires stays a two-wire diagonal.
        ops = (ops[0], ops[1], [w[:2] if nm == "MultiRZ" else w for nm, w in zip(ops[0], ops[2])],
               ops[3], ops[4])
        npar = sum(1 for p in ops[1] if p)
        k = int(rng.integers(1, npar + 1)) if npar else 0
        tp = sorted(int(x) for x in rng.choice(npar, size=k, replace=False)) if npar else [0]
        r = int(rng.integers(3))
        if r == 0:
            ob = random_hamiltonian(n, 4, rng, max_weight=min(3, n))
        elif r == 1:
            ob = ("named", ("PauliX", "PauliY", "PauliZ", "Hadamard")[int(rng.integers(4))],
                  [int(rng.integers(n))])
        else:
            ob = ("ham", [1.0] * n, [("named", "PauliZ", [w]) for w in range(n)])
        if npar == 0:
            ops = (ops[0] + ["RX"], ops[1] + [[0.3]], ops[2] + [[0]], ops[3] + [False],
                   ops[4] + [[]])
        circuits.append((n, ops, ob))
        tps.append(tp)
    return circuits, tps


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_batched_resident_random(capi, dtype):
    rng = np.random.default_rng(71)
    circuits, tps = _random_batch(rng, 40, 1, 8)
    _check(capi, circuits, tps, dtype)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("tile_bits", [6, 8])
def test_batched_streamed_random(capi, monkeypatch, dtype, tile_bits):
    """Tile width below n: circuits of up to 11 qubits go through global
    memory in 2^(n-T) tiles per pass; resident ones are mixed in."""
    monkeypatch.setenv("PLK_BATCH_TILE_BITS", str(tile_bits))
    rng = np.random.default_rng(72 + tile_bits)
    circuits, tps = _random_batch(rng, 24, 3, 11, nops=40)
    _check(capi, circuits, tps, dtype)


def test_batched_matches_per_circuit_engine(capi):
    rng = np.random.default_rng(73)
    circuits, tps = _random_batch(rng, 12, 4, 10)
    ev, rows = _batched(capi, circuits, tps, np.complex128)
    for i, (n, ops, ob) in enumerate(circuits):
        s = capi.State(n, np.complex128)
        jac = capi.adjoint_jacobian(s, [to_capi(capi, ob)], capi.Ops(*ops), tps[i], True)[0]
        np.testing.assert_allclose(rows[i], jac, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_batched_c4_sample(capi, dtype):
    """The 6-16 qubit layered workload (RY, RZ on every wire, CNOT chain;
    observable sum_i Z_i), every parameter trainable, 48 circuits."""
    circuits = c4_workload(count=48, seed=5)
    ev, rows = _check(capi, circuits, None, dtype)
    assert [len(r) for r in rows] == [sum(1 for p in c[1][1] if p) for c in circuits]


def test_batched_deterministic(capi):
    circuits = c4_workload(count=64, seed=9, nmax=14)
    a = _batched(capi, circuits, None, np.complex64)
    b = _batched(capi, circuits, None, np.complex64)
    np.testing.assert_array_equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        np.testing.assert_array_equal(x, y)


def test_batched_trainable_subsets_and_inverse(capi):
    """Reference bookkeeping: skipped parameters, inverse ops (sign flip of
    the generator term) and indices past the last parameter (zero columns)."""
    n = 5
    ops = (["RX", "CRY", "IsingXY", "RZ", "CNOT", "PhaseShift"],
           [[0.4], [1.1], [-0.7], [2.2], [], [0.9]],
           [[0], [1, 3], [2, 4], [3], [0, 2], [4]],
           [False, True, False, True, False, True], [[]] * 6)
    ob = ("ham", [0.5, -1.2], [("named", "PauliX", [1]),
                               ("tensor", [("named", "PauliY", [2]), ("named", "PauliZ", [4])])])
    tps = [[0, 2, 4], [1, 3], [0, 1, 2, 3, 4, 7]]
    circuits = [(n, ops, ob)] * 3
    ev, rows = _batched(capi, circuits, tps, np.complex128)
    for i, tp in enumerate(tps):
        e, jac = _oracle(n, ops, ob, tp)
        np.testing.assert_allclose(rows[i], jac, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(ev[i], e, rtol=1e-10, atol=1e-12)
    assert rows[2][5] == 0.0


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_batched_large_streamed_circuits(capi, dtype):
    """17-20 qubits at the default tile width: every pass streams 2^(n-T)
    tiles through LDS; mixed with small resident circuits."""
    rng = np.random.default_rng(81)
    circuits, tps = _random_batch(rng, 4, 17, 20, nops=50)
    small, stps = _random_batch(rng, 6, 4, 9, nops=20)
    _check(capi, circuits + small, tps + stps, dtype)


def _reparam(circuits, rng, special=None):
    """The same circuits with fresh angles (or `special` for every angle)."""
    out = []
    for n, ops, ob in circuits:
        params = [[float(special if special is not None else rng.uniform(0, 2 * np.pi))] * len(p)
                  if p else [] for p in ops[1]]
        out.append((n, (ops[0], params, ops[2], ops[3], ops[4]), ob))
    return out


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_batched_plan_reuse_new_angles(capi, dtype):
    """Repeated structure reuses the resident plan and only re-uploads the
    angles: each call must match the oracle for its own angles, including a
    first call at angle 0 (where RY / RX are diagonal) followed by generic
    angles, and a change of every other parameter or observable (replanned)."""
    rng = np.random.default_rng(7300)
    circuits, tps = _random_batch(rng, 24, 2, 9)
    _check(capi, _reparam(circuits, rng, special=0.0), tps, dtype)
    for _ in range(2):
        _check(capi, _reparam(circuits, rng), tps, dtype)
    # Observable coefficients changed: a different plan.
    changed = [(n, ops, ("ham", [0.5 * (k + 1) for k in range(n)], [("named", "PauliZ", [w]) for w in range(n)]))
               for n, ops, _ in circuits]
    _check(capi, changed, tps, dtype)
    _check(capi, _reparam(changed, rng), tps, dtype)
    # Trainable set changed.
    tps2 = [tp[:1] for tp in tps]
    _check(capi, _reparam(changed, rng), tps2, dtype)


def test_batched_plan_reuse_c4_sample(capi):
    """The C4 workload's structure (RY / RZ layers, CNOT chain) with two
    angle sets in a row, complex64 and complex128 interleaved."""
    rng = np.random.default_rng(7400)
    circuits = [(n, ops, ob) for n, ops, ob in c4_workload(count=24, seed=99)]
    for _ in range(2):
        fresh = _reparam(circuits, rng)
        _check(capi, fresh, None, np.complex64)
        _check(capi, fresh, None, np.complex128)
