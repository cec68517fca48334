"""Fused 5-qubit blocks on the matrix cores (plk_apply_ops_blocks, and
4-5-wire dense matrices on complex64 states) against the oracle.

complex64 blocks are FP32 products accumulated by v_mfma_f32_16x16x4f32,
so their bar is the complex64 one: 2e-5 absolute per amplitude for circuits
of up to a few hundred gates on normalised states, against the oracle's
complex128 result; complex128 blocks (v_mfma_f64_16x16x4f64) meet 1e-12. The large-state case compares the complex64 block path with the
engine's own complex128 path (the oracle's restatement cannot hold 2^28+
amplitudes in this test's time), on amplitudes, norm and the TFIM energy."""
import numpy as np
import pytest

import oracle as O
from helpers import c5_workload, random_ops, random_state, random_unitary, to_capi, to_oracle

pytestmark = pytest.mark.gpu

ONE_TWO_THREE = ["Hadamard", "RX", "RY", "RZ", "PhaseShift", "S", "T", "PauliX", "CNOT", "CZ", "CRX",
                 "SWAP", "IsingXX", "IsingZZ", "IsingXY", "Toffoli", "CSWAP", "SingleExcitation",
                 "DoubleExcitation", "MultiRZ"]


def _tol(dtype):
    # complex128 blocks run on FP64 MFMA: the 1e-10 bar of every FP64 result.
    return 2e-5 if dtype == np.complex64 else 1e-12


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("n", [9, 10, 12])
@pytest.mark.parametrize("seed", range(3))
def test_blocks_random_circuit(capi, n, seed, dtype):
    rng = np.random.default_rng(5000 + 10 * n + seed)
    ops = random_ops(n, 120, rng, names=ONE_TWO_THREE, with_matrix=(seed == 1))
    psi = random_state(n, rng)
    ref = O.apply_ops(psi.copy(), O.Ops(*ops))
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    nb = s.apply_ops_blocks(capi.Ops(*ops))
    assert 0 < nb <= len(ops[0])
    np.testing.assert_allclose(s.download(), ref, atol=_tol(dtype))


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("m", [4, 5])
@pytest.mark.parametrize("inverse", [False, True])
def test_dense_matrix_on_matrix_cores(capi, m, inverse, dtype):
    rng = np.random.default_rng(5100 + m)
    n = 11
    U = random_unitary(1 << m, rng)
    wires = [int(w) for w in rng.choice(n, size=m, replace=False)]
    psi = random_state(n, rng)
    ref = O.apply_matrix(psi.copy(), U, wires, inverse)
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    s.apply_matrix(U, wires, inverse)
    np.testing.assert_allclose(s.download(), ref, atol=_tol(dtype))


def test_blocks_c5_structure_vs_oracle(capi):
    """The C5 circuit (Haar SU(2) layers, CZ bricks, TFIM) at 14 qubits."""
    n, ops, ham = c5_workload(n=14, layers=12, seed=3)
    psi = O.zero_state(n)
    ref = O.apply_ops(psi.copy(), O.Ops(*ops))
    e_ref = O.expval(ref, to_oracle(ham))
    s = capi.State(n, np.complex64)
    nb = s.apply_ops_blocks(capi.Ops(*ops))
    print(f"c5 14q: {len(ops[0])} ops in {nb} blocks", flush=True)
    np.testing.assert_allclose(s.download(), ref, atol=2e-5)
    assert abs(s.expval(to_capi(capi, ham)) - e_ref) < 1e-4


def test_blocks_large_state_vs_complex128_path(capi):
    """24 qubits, 8 layers: complex64 blocks vs the complex128 plain path."""
    n, ops, ham = c5_workload(n=24, layers=8, seed=4)
    a = capi.State(n, np.complex64)
    nb = a.apply_ops_blocks(capi.Ops(*ops))
    b = capi.State(n, np.complex128)
    b.apply_ops(capi.Ops(*ops))
    print(f"c5 24q: {len(ops[0])} ops in {nb} blocks", flush=True)
    h = to_capi(capi, ham)
    assert abs(a.expval(h) - b.expval(h)) < 1e-4
    idx = np.random.default_rng(6).integers(0, 1 << n, size=4096)
    np.testing.assert_allclose(a.download()[idx], b.download()[idx], atol=2e-5)
    assert abs(np.linalg.norm(a.download()) - 1.0) < 1e-4


def test_blocks_reject(capi):
    ops = capi.Ops(["RX"], [[0.1]], [[0]], [False], [[]])
    s8 = capi.State(8, np.complex64)
    with pytest.raises(capi.PLException, match="9 qubits"):
        s8.apply_ops_blocks(ops)
