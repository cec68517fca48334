"""plk_expval_obs_noalloc against plk_expval_obs and the oracle.

The allocation-free entry point keeps its partial sums and term tables in a
fixed device array shared by every call on the GPU, so the test makes two
calls in a row, the second with a smaller state and fewer Pauli words: stale
partials or tables left by the first call would show in the second. Parity
bar as for every FP64 expectation value: 1e-12 relative + 1e-13 absolute
against the oracle, 1e-5 absolute for complex64 states; against
plk_expval_obs (same kernels, same grid) the results agree to 1e-14.
"""
import numpy as np
import pytest

import oracle as O
from helpers import random_hamiltonian, random_state, to_capi, to_oracle

pytestmark = pytest.mark.gpu


def _tensor_ham(n, rng, terms):
    """Hamiltonian mixing Pauli words, a Hadamard factor and an Identity term."""
    ham = random_hamiltonian(n, terms, rng)
    coeffs = list(ham[1]) + [0.7, -0.3]
    words = list(ham[2]) + [("tensor", [("named", "Hadamard", [0]), ("named", "PauliY", [n - 1])]),
                            ("named", "Identity", [1])]
    return ("ham", coeffs, words)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_noalloc_matches_expval_obs_twice(capi, dtype):
    rng = np.random.default_rng(4242)
    cases = [(18, 700), (9, 5)]  # second call: smaller state, fewer words
    for n, terms in cases:
        psi = random_state(n, rng)
        ob = _tensor_ham(n, rng, terms)
        s = capi.State(n, dtype)
        s.upload(psi.astype(dtype))
        h = to_capi(capi, ob)
        got = s.expval_noalloc(h)
        base = s.expval(h)
        ref = O.expval(psi, to_oracle(ob))
        np.testing.assert_allclose(got, base, rtol=1e-14, atol=1e-14)
        if dtype == np.complex128:
            np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13)
        else:
            np.testing.assert_allclose(got, ref, atol=1e-5 * max(1.0, np.sum(np.abs(ob[1]))))


def test_noalloc_single_words_and_empty_sum(capi):
    rng = np.random.default_rng(4243)
    n = 11
    psi = random_state(n, rng)
    s = capi.State(n)
    s.upload(psi)
    for ob in [("named", "PauliZ", [3]), ("named", "PauliX", [0]), ("named", "Identity", [2]),
               ("tensor", [("named", "PauliY", [4]), ("named", "PauliX", [7])])]:
        h = to_capi(capi, ob)
        np.testing.assert_allclose(s.expval_noalloc(h), O.expval(psi, to_oracle(ob)),
                                   rtol=1e-12, atol=1e-13)
    empty = capi.Obs.hamiltonian([], [])
    assert s.expval_noalloc(empty) == 0.0


def test_noalloc_limits(capi):
    rng = np.random.default_rng(4244)
    n = 12
    s = capi.State(n)
    s.upload(random_state(n, rng))
    big = to_capi(capi, random_hamiltonian(n, 4097, rng))
    with pytest.raises(capi.PLException, match="at most 4096"):
        s.expval_noalloc(big)
    herm = capi.Obs.hermitian(np.eye(2), [0])
    with pytest.raises(capi.PLException, match="not a sum"):
        s.expval_noalloc(herm)
    # The array is left usable after a rejected call.
    ob = random_hamiltonian(n, 40, rng)
    psi = s.download()
    np.testing.assert_allclose(s.expval_noalloc(to_capi(capi, ob)), O.expval(psi, to_oracle(ob)),
                               rtol=1e-12, atol=1e-13)
