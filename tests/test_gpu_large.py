"""Large-size cases checked through size-independent properties.

- A 26-qubit (1 GiB) stream of 2000 random gates followed by its
  inverse returns to |0...0>, and the norm after the forward stream is 1.
- A 29-qubit complex128 state (8 GiB) exercises the kernels' 64-bit byte
  offsets (states above 4 GiB): same round trip.
- A 64-word Pauli Hamiltonian at 16 qubits: the Pauli-word kernel, the CSR
  kernel and the oracle agree."""
import numpy as np
import pytest

import oracle as O
from helpers import dense_pauli_sparse, random_hamiltonian, random_ops, random_state, to_capi, to_oracle

pytestmark = pytest.mark.gpu

C2_GATES = ["Hadamard", "RX", "RY", "RZ", "CNOT", "CZ", "CRX", "SWAP", "IsingXX", "IsingZZ"]


def _roundtrip(capi, n, count, dtype, seed):
    rng = np.random.default_rng(seed)
    names, params, wires, inv, mats = random_ops(n, count, rng, names=C2_GATES, allow_inverse=False)
    s = capi.State(n, dtype)
    s.apply_ops(capi.Ops(names, params, wires, inv, mats))
    norm = s.inner(s)
    rev = (names[::-1], params[::-1], wires[::-1], [True] * count, mats[::-1])
    s.apply_ops(capi.Ops(*rev))
    return s, norm


def test_c2_26q_stream_roundtrip(capi):
    s, norm = _roundtrip(capi, 26, 2000, np.complex128, 2026)
    assert abs(norm - 1.0) < 1e-10
    z = capi.State(26)
    amp0 = z.inner(s)
    assert abs(amp0 - 1.0) < 1e-9, amp0


def test_state_above_4gib_roundtrip(capi):
    """2^29 complex128 amplitudes = 8 GiB: the 64-bit offset kernels."""
    s, norm = _roundtrip(capi, 29, 120, np.complex128, 29)
    assert abs(norm - 1.0) < 1e-10
    assert abs(s.probs([0, 13, 28])[0] - 1.0) < 1e-9


def test_pauli_csr_oracle_agree_16q(capi):
    n = 16
    rng = np.random.default_rng(64)
    ham = random_hamiltonian(n, 64, rng)
    psi = random_state(n, rng)
    s = capi.State(n)
    s.upload(psi)
    e_pauli = s.expval(to_capi(capi, ham))
    d, i, p = dense_pauli_sparse(n, ham)
    e_csr = s.expval(capi.Obs.sparse(d, i, p, list(range(n))))
    e_ref = O.expval(psi, to_oracle(ham))
    assert abs(e_pauli - e_ref) < 1e-10 * max(1.0, abs(e_ref))
    assert abs(e_csr - e_ref) < 1e-10 * max(1.0, abs(e_ref))
