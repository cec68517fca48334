"""GPU parity: the HIP engine through its C ABI against the CPU oracle.

Tolerances: complex128 states match the oracle to 1e-12 absolute per
amplitude / 1e-10 relative per Jacobian entry; complex64 states (FP64
reductions) to 2e-5 absolute.
"""
import numpy as np
import pytest

import oracle as O
from helpers import (GATES, GEN_GATES, c1_workload, dense_pauli_sparse, gate_wires,
                     random_hamiltonian, random_hermitian, random_ops, random_state,
                     random_unitary, to_capi, to_oracle)

pytestmark = pytest.mark.gpu

C128_ATOL = 1e-12
C64_ATOL = 2e-5


def _tol(dtype):
    return C128_ATOL if dtype == np.complex128 else C64_ATOL


def _gpu_state(capi, psi, dtype):
    n = int(np.log2(psi.size))
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    return s


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_state_roundtrip_and_basis(capi, dtype):
    rng = np.random.default_rng(1)
    psi = random_state(10, rng)
    s = _gpu_state(capi, psi, dtype)
    np.testing.assert_allclose(s.download(), psi.astype(dtype), atol=0)
    s.reset()
    out = s.download()
    assert out[0] == 1 and np.count_nonzero(out) == 1
    s.set_basis(37)
    out = s.download()
    assert out[37] == 1 and np.count_nonzero(out) == 1
    s.set_values([3, 5, 700], np.array([0.6, 0.8j, 0.0]))
    out = s.download()
    assert out[3] == pytest.approx(0.6) and out[5] == pytest.approx(0.8j)
    assert np.count_nonzero(out) == 2


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("name", [g for g in GATES if g != "Identity"])
def test_named_gate(capi, name, dtype):
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    for n in (4, 9):
        nw, npar = GATES[name]
        if nw > n:
            continue
        for trial in range(3):
            psi = random_state(n, rng)
            wires = gate_wires(name, n, rng, k_multi=min(n, 4))
            params = list(rng.uniform(-np.pi, 2 * np.pi, size=npar))
            inv = bool(trial % 2)
            ref = O.apply(psi.copy(), name, wires, inv, params)
            s = _gpu_state(capi, psi, dtype)
            s.apply(name, wires, inv, params)
            np.testing.assert_allclose(s.download(), ref, atol=_tol(dtype),
                                       err_msg=f"{name} wires={wires} inv={inv} n={n}")


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_matrix_op(capi, k, dtype):
    rng = np.random.default_rng(100 + k)
    n = 8
    for inv in (False, True):
        psi = random_state(n, rng)
        U = random_unitary(2**k, rng)
        wires = [int(x) for x in rng.choice(n, size=k, replace=False)]
        ref = O.apply_matrix(psi.copy(), U, wires, inv)
        s = _gpu_state(capi, psi, dtype)
        s.apply_matrix(U, wires, inv)
        np.testing.assert_allclose(s.download(), ref, atol=_tol(dtype))


@pytest.mark.parametrize("name", GEN_GATES)
def test_generator(capi, name):
    rng = np.random.default_rng(7)
    n = 6
    psi = random_state(n, rng)
    wires = gate_wires(name, n, rng)
    ref = psi.copy()
    s_ref = O.apply_generator(ref, name, wires)
    s = _gpu_state(capi, psi, np.complex128)
    sc = s.apply_generator(name, wires)
    assert sc == s_ref
    np.testing.assert_allclose(s.download(), ref, atol=C128_ATOL)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_apply_ops_fused_matches_sequential(capi, dtype):
    rng = np.random.default_rng(11)
    n = 10
    names, params, wires, inv, mats = random_ops(n, 80, rng, with_matrix=True)
    psi = random_state(n, rng)
    ref = O.apply_ops(psi.copy(), O.Ops(names, params, wires, inv, mats))
    s = _gpu_state(capi, psi, dtype)
    s.apply_ops(capi.Ops(names, params, wires, inv, mats))
    np.testing.assert_allclose(s.download(), ref, atol=50 * _tol(dtype))


def _expval_cases(n, rng):
    cases = [("named", "PauliX", [1]), ("named", "PauliY", [n - 1]), ("named", "PauliZ", [0]),
             ("named", "Hadamard", [2]), ("named", "Identity", [0]),
             ("tensor", [("named", "PauliX", [0]), ("named", "PauliY", [2]),
                         ("named", "PauliZ", [n - 1])]),
             ("tensor", [("named", "Hadamard", [1]), ("named", "PauliZ", [3])]),
             random_hamiltonian(n, 12, rng),
             ("herm", random_hermitian(4, rng), [3, 1]),
             ("herm", random_hermitian(8, rng), [0, 4, 2]),
             ("tensor", [("herm", random_hermitian(2, rng), [0]), ("named", "PauliX", [1])]),
             ("ham", [0.3, -1.1], [("herm", random_hermitian(4, rng), [1, 2]),
                                   ("named", "PauliZ", [0])])]
    ham = random_hamiltonian(n, 6, rng)
    d, i, p = dense_pauli_sparse(n, ham)
    cases.append(("sparse", d, i, p, list(range(n))))
    return cases


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_expval_observables(capi, dtype):
    rng = np.random.default_rng(21)
    n = 7
    psi = random_state(n, rng)
    s = _gpu_state(capi, psi, dtype)
    for ob in _expval_cases(n, rng):
        ref = O.expval(psi, to_oracle(ob))
        got = s.expval(to_capi(capi, ob))
        assert got == pytest.approx(ref, abs=1e3 * _tol(dtype)), ob[0]


def test_expval_named_matrix_csr(capi):
    rng = np.random.default_rng(22)
    n = 8
    psi = random_state(n, rng)
    s = _gpu_state(capi, psi, np.complex128)
    for name in ("PauliX", "PauliY", "PauliZ", "Hadamard", "Identity"):
        ref = O.expval(psi, O.Named(name, [5]))
        assert s.expval_named(name, [5]) == pytest.approx(ref, abs=1e-12)
    H = random_hermitian(4, rng)
    ref = O.expval(psi, O.Hermitian(H, [6, 2]))
    assert s.expval_matrix(H, [6, 2]) == pytest.approx(ref, abs=1e-12)
    d, i, p = dense_pauli_sparse(n, random_hamiltonian(n, 5, rng))
    assert s.expval_csr(d, i, p) == pytest.approx(O.expval_csr(psi, d, i, p), abs=1e-12)


def test_inner_product(capi):
    rng = np.random.default_rng(23)
    a, b = random_state(9, rng), random_state(9, rng)
    sa, sb = _gpu_state(capi, a, np.complex128), _gpu_state(capi, b, np.complex128)
    assert sa.inner(sb) == pytest.approx(np.vdot(a, b), abs=1e-13)


def _adjoint_pair(capi, psi, ops, obs, tp, apply_ops, dtype):
    names, params, wires, inv, mats = ops
    ref = O.adjoint_jacobian(psi, [to_oracle(o) for o in obs], O.Ops(names, params, wires, inv, mats),
                             tp, apply_ops)
    s = _gpu_state(capi, psi, dtype)
    got = capi.adjoint_jacobian(s, [to_capi(capi, o) for o in obs],
                                capi.Ops(names, params, wires, inv, mats), tp, apply_ops)
    return ref, got


def _assert_jac(ref, got, dtype):
    if dtype == np.complex128:
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)
    else:
        np.testing.assert_allclose(got, ref, atol=1e-5)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("seed", range(6))
def test_adjoint_random_circuits(capi, seed, dtype):
    rng = np.random.default_rng(1000 + seed)
    n = 6 + seed % 3
    ops = random_ops(n, 40, rng, with_matrix=(seed % 2 == 1))
    npar = sum(1 for p in ops[1] if p)
    tp = sorted(int(x) for x in rng.choice(npar, size=max(1, npar - seed), replace=False))
    nobs = [1, 3, 9, 5, 2, 12][seed]
    obs = []
    for j in range(nobs):
        r = j % 4
        if r == 0:
            obs.append(("named", ("PauliX", "PauliY", "PauliZ", "Hadamard")[j % 4], [j % n]))
        elif r == 1:
            obs.append(random_hamiltonian(n, 5, rng))
        elif r == 2:
            obs.append(("tensor", [("named", "PauliZ", [0]), ("named", "PauliX", [n - 1])]))
        else:
            obs.append(("herm", random_hermitian(4, rng), [1, 3]))
    psi = random_state(n, rng)
    ref, got = _adjoint_pair(capi, psi, ops, obs, tp, seed % 2 == 0, dtype)
    _assert_jac(ref, got, dtype)


def test_adjoint_reference_goldens(capi):
    """Goldens of src/tests/Test_AdjointDiffKokkos.cpp (180-390)."""
    p = [-np.pi / 7, np.pi / 5, 2 * np.pi / 3]
    z3 = np.zeros(8, complex)
    z3[0] = 1
    ops = (["RX"] * 3, [[p[0]], [p[1]], [p[2]]], [[0], [1], [2]], [False] * 3, [[]] * 3)
    obs = [("tensor", [("named", "PauliZ", [0]), ("named", "PauliZ", [1]), ("named", "PauliZ", [2])])]
    _, got = _adjoint_pair(capi, z3, ops, obs, [0, 1, 2], True, np.complex128)
    np.testing.assert_allclose(got[0], [-0.1755096592645253, 0.26478810666384334, -0.6312451595102775],
                               atol=1e-7)
    ops = (["RZ", "RY", "RZ", "CNOT", "CNOT", "RZ", "RY", "RZ"],
           [[p[0]], [p[1]], [p[2]], [], [], [p[0]], [p[1]], [p[2]]],
           [[0], [0], [0], [0, 1], [1, 2], [1], [1], [1]], [False] * 8, [[]] * 8)
    obs = [("tensor", [("named", "PauliX", [0]), ("named", "PauliX", [1]), ("named", "PauliX", [2])])]
    _, got = _adjoint_pair(capi, z3, ops, obs, [0, 1, 2, 3, 4, 5], True, np.complex128)
    np.testing.assert_allclose(got[0], [0, -0.674214427, 0.275139672, 0.275139672, -0.0129093062,
                                        0.323846156], atol=1e-7)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_adjoint_c1_workload(capi, dtype):
    """12 qubits, 144 parameters, 12 PauliZ observables (complex128 at 1e-10)."""
    n, ops, obs = c1_workload()
    psi = O.zero_state(n)
    tp = list(range(144))
    ref, got = _adjoint_pair(capi, psi, ops, obs, tp, True, dtype)
    _assert_jac(ref, got, dtype)


def test_adjoint_errors(capi):
    s = capi.State(3)
    ops = capi.Ops(["RX"], [[0.1]], [[0]], [False])
    with pytest.raises(capi.PLException, match="No trainable parameters provided."):
        capi.adjoint_jacobian(s, [capi.Obs.named("PauliZ", [0])], ops, [])
    ops = capi.Ops(["Rot"], [[0.1, 0.2, 0.3]], [[0]], [False])
    with pytest.raises(capi.PLException, match="not supported using the adjoint"):
        capi.adjoint_jacobian(s, [capi.Obs.named("PauliZ", [0])], ops, [0])
    with pytest.raises(capi.PLException, match="disjoint"):
        capi.Obs.tensor([capi.Obs.named("PauliZ", [0]), capi.Obs.named("PauliX", [0])])
    with pytest.raises(capi.PLException, match="Generator does not exist"):
        s.apply_generator("CNOT", [0, 1])


# ----------------------------------------------------------------------------
# Fused passes at 14+ qubits: register-tile passes (the default) and LDS tile
# passes (PLK_RTILE=0); PLK_RTILE=0 with PLK_TILES=0 forces register
# segments for cross-checks.
# ----------------------------------------------------------------------------
def _layered_ops(n, layers, rng):
    names, params, wires = [], [], []
    for layer in range(layers):
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
    return names, params, wires, [False] * len(names), [[] for _ in names]


@pytest.fixture(params=["rtile", "lds"])
def engine(request, monkeypatch):
    monkeypatch.setenv("PLK_RTILE", "1" if request.param == "rtile" else "0")
    return request.param


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [14, 16, 18])
def test_adjoint_tiles_layered(capi, engine, n, dtype):
    rng = np.random.default_rng(300 + n)
    ops = _layered_ops(n, 3, rng)
    obs = [("named", "PauliZ", [0]), ("named", "PauliX", [n - 1]), random_hamiltonian(n, 6, rng),
           ("tensor", [("named", "PauliY", [2]), ("named", "PauliZ", [n - 2])]), ("named", "PauliZ", [n // 2])]
    npar = sum(1 for p in ops[1] if p)
    ref, got = _adjoint_pair(capi, O.zero_state(n), ops, obs, list(range(npar)), True, dtype)
    _assert_jac(ref, got, dtype)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("seed", range(3))
def test_adjoint_tiles_random(capi, engine, seed, dtype):
    rng = np.random.default_rng(400 + seed)
    n = [15, 17, 18][seed]
    ops = random_ops(n, 120, rng, with_matrix=(seed != 1))
    npar = sum(1 for p in ops[1] if p)
    tp = sorted(int(x) for x in rng.choice(npar, size=npar - 3 * seed, replace=False))
    obs = [("named", "PauliZ", [j % n]) for j in range(2 + 3 * seed)] + [random_hamiltonian(n, 5, rng)]
    psi = random_state(n, rng)
    ref, got = _adjoint_pair(capi, psi, ops, obs, tp, seed != 2, dtype)
    _assert_jac(ref, got, dtype)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [20, 21])
def test_apply_ops_tiles(capi, engine, n, dtype):
    rng = np.random.default_rng(500 + n)
    names, params, wires, inv, mats = random_ops(n, 150, rng, with_matrix=True)
    psi = random_state(n, rng)
    ref = O.apply_ops(psi.copy(), O.Ops(names, params, wires, inv, mats))
    s = _gpu_state(capi, psi, dtype)
    s.apply_ops(capi.Ops(names, params, wires, inv, mats))
    np.testing.assert_allclose(s.download(), ref, atol=50 * _tol(dtype))


def test_tiles_match_segments(capi, monkeypatch):
    rng = np.random.default_rng(600)
    n = 18
    ops = _layered_ops(n, 2, rng)
    obs = [("named", "PauliZ", [3]), random_hamiltonian(n, 4, rng)]
    npar = sum(1 for p in ops[1] if p)
    s = _gpu_state(capi, O.zero_state(n), np.complex128)
    cobs = [to_capi(capi, o) for o in obs]
    cops = capi.Ops(*ops)
    rtiled = capi.adjoint_jacobian(s, cobs, cops, list(range(npar)), True)
    monkeypatch.setenv("PLK_RTILE", "0")
    tiled = capi.adjoint_jacobian(s, cobs, cops, list(range(npar)), True)
    monkeypatch.setenv("PLK_TILES", "0")
    segs = capi.adjoint_jacobian(s, cobs, cops, list(range(npar)), True)
    np.testing.assert_allclose(tiled, segs, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(rtiled, segs, rtol=1e-11, atol=1e-13)
