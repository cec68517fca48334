"""Dense matrices on up to 12 wires (expectation values without a state
copy, applications out of place) and the device's dense-Hamiltonian route
(lightning_kokkos.py:521-534), against the oracle. complex128: 1e-10
relative + 1e-12 absolute; complex64 (FP32 matrix entries, FP64 sums): 1e-4
relative to the matrix scale."""
import numpy as np
import pytest

import oracle as O
from helpers import random_hermitian, random_state, random_unitary, to_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,m", [(10, 3), (12, 9), (13, 12), (12, 12)])
def test_expval_dense_matrix(capi, n, m, dtype):
    rng = np.random.default_rng(9000 + 13 * n + m)
    H = random_hermitian(1 << m, rng) / np.sqrt(1 << m)
    wires = [int(w) for w in rng.choice(n, size=m, replace=False)]
    psi = random_state(n, rng)
    ref = O.expval(psi, O.Hermitian(H, wires))
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    got = s.expval_matrix(H, wires)
    tol = 1e-10 * max(1.0, abs(ref)) if dtype == np.complex128 else 1e-4
    assert abs(got - ref) < tol, (got, ref)
    # The observable handle takes the same path.
    assert abs(s.expval(capi.Obs.hermitian(H, wires)) - ref) < tol


@pytest.mark.parametrize("m", [9, 10])
def test_apply_dense_matrix_beyond_eight_wires(capi, m):
    rng = np.random.default_rng(9100 + m)
    n = 12
    U = random_unitary(1 << m, rng)
    wires = [int(w) for w in rng.choice(n, size=m, replace=False)]
    psi = random_state(n, rng)
    ref = O.apply_matrix(psi.copy(), U, wires, False)
    s = capi.State(n)
    s.upload(psi)
    s.apply_matrix(U, wires, False)
    np.testing.assert_allclose(s.download(), ref, atol=1e-12)


def test_device_dense_hamiltonian_route(capi):
    from pennylane_lightning_kokkos_amd.device import LightningKokkosAMD

    rng = np.random.default_rng(9200)
    n = 11
    dev = LightningKokkosAMD(n)
    psi = random_state(n, rng)
    dev.syncH2D(psi)
    H4 = random_hermitian(16, rng)
    ham = ("ham", [0.7, -0.4, 1.1],
           [("herm", H4, [3, 8, 1, 10]),
            ("tensor", [("named", "PauliX", [0]), ("named", "PauliZ", [5])]),
            ("named", "Hadamard", [7])])
    ref = O.expval(psi, to_oracle(ham))
    got = dev.expval(ham)
    assert abs(got - ref) < 1e-10 * max(1.0, abs(ref))
