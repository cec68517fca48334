"""The CPU oracle against golden vectors of the reference's own tests.

Fixtures: tests/golden/reference_goldens.json (transcribed from
pennylane_lightning_kokkos/src/tests/*.cpp by transcribe_reference_goldens.py)
and the adjoint goldens of src/tests/Test_AdjointDiffKokkos.cpp.
Tolerances follow the reference's Catch2 Approx (relative 1.2e-5) and the
6-digit precision of some expected vectors.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import to_oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))["goldens"]


def _c(pairs):
    a = np.asarray(pairs, dtype=np.float64)
    return np.ascontiguousarray(a[:, 0] + 1j * a[:, 1])


def _gid(g):
    if g["kind"] == "gate_seq":
        return "seq-" + "-".join(f"{o[0]}{o[1]}" for o in g["ops"])
    return f"{g['gate']}-{g['wires']}-{g['params']}-inv{g['inverse']}"


@pytest.mark.parametrize("g", [g for g in GOLD if g["kind"] in ("gate", "gate_seq")], ids=_gid)
def test_gate_goldens(g):
    psi = _c(g["init"])
    ops = g["ops"] if g["kind"] == "gate_seq" else [[g["gate"], g["wires"], g["params"], g["inverse"]]]
    for name, wires, params, inv in ops:
        O.apply(psi, name, wires, inv, params)
    # Entries the reference test checks ("check"), or the whole vector.
    idx = g.get("check", list(range(psi.size)))
    np.testing.assert_allclose(psi[idx], _c(g["expected"])[idx], atol=2e-6, rtol=1.2e-5, err_msg=g["source"])


def test_golden_coverage_of_gate_kernels():
    """Every named gate of the reference test files NonParam / Param has at
    least one transcribed golden."""
    have = {g["gate"] for g in GOLD if g["kind"] == "gate"}
    have |= {o[0] for g in GOLD if g["kind"] == "gate_seq" for o in g["ops"]}
    want = {"Hadamard", "PauliX", "PauliY", "PauliZ", "S", "T", "CNOT", "SWAP", "CZ", "Toffoli", "CSWAP",
            "RX", "RY", "RZ", "PhaseShift", "ControlledPhaseShift", "Rot", "CRot", "IsingXX", "IsingXY",
            "IsingYY", "IsingZZ", "MultiRZ", "SingleExcitation", "SingleExcitationMinus",
            "SingleExcitationPlus", "DoubleExcitation", "DoubleExcitationMinus", "DoubleExcitationPlus"}
    assert want <= have, sorted(want - have)


def test_expval_goldens():
    for g in GOLD:
        if g["kind"] == "expval_matrix":
            psi = _c(g["init"])
            M = _c(g["matrix"]).reshape(8, 8)
            val = O.expval(psi, O.Hermitian(M, g["wires"]))
            assert val == pytest.approx(g["expected"], rel=g["rel"], abs=1e-9), g["source"]
        elif g["kind"] == "expval_csr":
            psi = _c(g["init"])
            val = O.expval_csr(psi, _c(g["data"]), g["indices"], g["indptr"])
            assert val == pytest.approx(g["expected"], rel=g["rel"]), g["source"]
        elif g["kind"] == "expval_obs":
            psi = _c(g["init"])
            val = O.expval(psi, to_oracle(g["obs"]))
            assert val == pytest.approx(g["expected"], rel=g["rel"]), g["source"]


def test_linear_algebra_goldens():
    for g in GOLD:
        if g["kind"] == "spmv":
            y = O.spmv(_c(g["x"]), _c(g["data"]), g["indices"], g["indptr"])
            np.testing.assert_allclose(y, _c(g["expected"]), atol=1e-12, err_msg=g["source"])
        elif g["kind"] == "axpy":
            # y += alpha x (LinearAlgebraKokkos.hpp:30-61), a closed form.
            alpha = complex(*g["alpha"])
            np.testing.assert_allclose(_c(g["y"]) + alpha * _c(g["x"]), _c(g["expected"]), atol=1e-12)


P3 = [-np.pi / 7, np.pi / 5, 2 * np.pi / 3]


def _z(n):
    return O.zero_state(n)


@pytest.mark.parametrize("p", P3)
def test_adjoint_closed_forms(p):
    """Test_AdjointDiffKokkos.cpp:34-178."""
    ops = O.Ops(["RX"], [[p]], [[0]], [False])
    assert O.adjoint_jacobian(_z(1), [O.Named("PauliZ", [0])], ops, [0], True)[0, 0] == pytest.approx(-np.sin(p))
    ops = O.Ops(["RY"], [[p]], [[0]], [False])
    assert O.adjoint_jacobian(_z(1), [O.Named("PauliX", [0])], ops, [0], True)[0, 0] == pytest.approx(np.cos(p))
    jac = O.adjoint_jacobian(_z(2), [O.Named("PauliZ", [0]), O.Named("PauliZ", [1])],
                             O.Ops(["RX"], [[p]], [[0]], [False]), [0], True)
    np.testing.assert_allclose(jac[:, 0], [-np.sin(p), 0.0], atol=1e-12)


def test_adjoint_goldens():
    """Test_AdjointDiffKokkos.cpp:131-455."""
    ops = O.Ops(["RX"] * 3, [[x] for x in P3], [[0], [1], [2]], [False] * 3)
    z3 = [O.Named("PauliZ", [w]) for w in range(3)]
    jac = O.adjoint_jacobian(_z(3), z3, ops, [0, 1, 2], True)
    np.testing.assert_allclose(np.diag(jac), -np.sin(P3), atol=1e-7)
    jac = O.adjoint_jacobian(_z(3), z3, ops, [0, 2], True)
    np.testing.assert_allclose([jac[0, 0], jac[1, 1], jac[2, 1]], [-np.sin(P3[0]), 0, -np.sin(P3[2])], atol=1e-7)
    zzz = O.Tensor(z3)
    jac = O.adjoint_jacobian(_z(3), [zzz], ops, [0, 1, 2], True)
    np.testing.assert_allclose(jac[0], [-0.1755096592645253, 0.26478810666384334, -0.6312451595102775], atol=1e-7)
    ops = O.Ops(["RZ", "RY", "RZ", "CNOT", "CNOT", "RZ", "RY", "RZ"],
                [[P3[0]], [P3[1]], [P3[2]], [], [], [P3[0]], [P3[1]], [P3[2]]],
                [[0], [0], [0], [0, 1], [1, 2], [1], [1], [1]], [False] * 8)
    xxx = O.Tensor([O.Named("PauliX", [w]) for w in range(3)])
    jac = O.adjoint_jacobian(_z(3), [xxx], ops, list(range(6)), True)
    np.testing.assert_allclose(jac[0], [0, -0.674214427, 0.275139672, 0.275139672, -0.0129093062, 0.323846156],
                               atol=1e-7)
    lp = [0.543, 0.54, 0.1, 0.5, 1.3, -2.3, 0.5, -0.5, 0.5]
    ops = O.Ops(["Hadamard", "RX", "CNOT", "RZ", "RY", "RZ", "RZ", "RY", "RZ", "RZ", "RY", "CNOT"],
                [[], [lp[0]], [], [lp[1]], [lp[2]], [lp[3]], [lp[4]], [lp[5]], [lp[6]], [lp[7]], [lp[8]], []],
                [[0], [0], [0, 1], [0], [0], [0], [0], [0], [0], [0], [1], [0, 1]], [False] * 12)
    jac = O.adjoint_jacobian(_z(2), [O.Tensor([O.Named("PauliX", [0]), O.Named("PauliZ", [1])])], ops,
                             [1, 2, 3], True)
    np.testing.assert_allclose(jac[0], [-0.71429188, 0.04998561, -0.71904837], rtol=1.2e-5)
    ham = O.Hamiltonian([0.47, 0.32, 0.96], z3)
    jac = O.adjoint_jacobian(_z(3), [ham], O.Ops(["RX"] * 3, [[x] for x in P3], [[0], [1], [2]], [False] * 3),
                             [0, 2], True)
    np.testing.assert_allclose(jac[0], [-0.47 * np.sin(P3[0]), -0.96 * np.sin(P3[2])], atol=1e-7)


def test_adjoint_decomposed_rot_goldens():
    """Test_AdjointDiffKokkos.cpp:262-317: RZ-RY-RZ on (|0> - |1>)/sqrt2, <Z>."""
    thetas = np.linspace(-2 * np.pi, 2 * np.pi, 7)
    expected = [[0, -9.90819496e-01, 0], [-8.18996553e-01, 1.62526544e-01, 0], [-0.203949, 0.48593716, 0],
                [0, 1, 0], [-2.03948985e-01, 4.85937177e-01, 0], [-8.18996598e-01, 1.62526487e-01, 0],
                [0, -9.90819511e-01, 0]]
    for th, ex in zip(thetas, expected):
        psi = np.array([1, -1], dtype=complex) / np.sqrt(2)
        ops = O.Ops(["RZ", "RY", "RZ"], [[th], [th**3], [np.sqrt(2) * th]], [[0]] * 3, [False] * 3)
        jac = O.adjoint_jacobian(psi, [O.Named("PauliZ", [0])], ops, [0, 1, 2], True)
        np.testing.assert_allclose(jac[0], ex, atol=1e-6)


def test_adjoint_hermitian_equals_tensor():
    """Test_AdjointDiffKokkos.cpp:457-492."""
    ops = O.Ops(["RX"] * 3, [[x] for x in P3], [[0], [1], [2]], [False] * 3)
    zz = O.Tensor([O.Named("PauliZ", [0]), O.Named("PauliZ", [1])])
    herm = O.Hermitian(np.diag([1, -1, -1, 1]).astype(complex), [0, 1])
    j1 = O.adjoint_jacobian(_z(3), [zz], ops, [0, 2], True)
    j2 = O.adjoint_jacobian(_z(3), [herm], ops, [0, 2], True)
    np.testing.assert_allclose(j1, j2, atol=1e-12)


def test_adjoint_error_messages():
    """AdjointDiffKokkos.hpp:410-411, 444-446; StateVectorKokkos.hpp:692-693."""
    with pytest.raises(O.OracleError, match="No trainable parameters provided."):
        O.adjoint_jacobian(_z(1), [O.Named("PauliZ", [0])], O.Ops(["RX"], [[0.1]], [[0]], [False]), [], True)
    with pytest.raises(O.OracleError, match="not supported using the adjoint"):
        O.adjoint_jacobian(_z(1), [O.Named("PauliZ", [0])], O.Ops(["Rot"], [[0.1, 0.2, 0.3]], [[0]], [False]),
                           [0], True)
    with pytest.raises(O.OracleError, match="Generator does not exist for CNOT"):
        O.apply_generator(_z(2), "CNOT", [0, 1])
