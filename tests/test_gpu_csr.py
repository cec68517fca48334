"""CSR expectation values and SpMV against the oracle.

The kernels pick 4, 8, 16, 32 or 64 lanes per row from the average row
length (each lane loads four entries per round), take
int32 or int64 indices as handed over (int32 is the complex64 serialisation,
_serialize.py:136-154) and check every row range and column index on the
device. These tests cover each lane width, both index widths, both
precisions, empty rows, the device-resident copy of a sparse observable
(expectation value, application, adjoint sweep) and malformed matrices.
Tolerances: complex128 1e-12 absolute (values of order 1), complex64 1e-5.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from helpers import random_ops, random_state

pytestmark = pytest.mark.gpu


def _random_csr(n, avg, rng, empty_rows=True):
    N = 1 << n
    m = sp.random(N, N, density=min(1.0, avg / N), random_state=np.random.RandomState(int(rng.integers(1 << 30))),
                  format="csr", dtype=np.float64)
    m = m + 1j * sp.random(N, N, density=min(1.0, avg / N),
                           random_state=np.random.RandomState(int(rng.integers(1 << 30))), format="csr")
    m = m.tocsr()
    if empty_rows:
        m = m.tolil()
        for r in rng.choice(N, size=N // 8, replace=False):
            m.rows[r] = []
            m.data[r] = []
        m = m.tocsr()
    m.sort_indices()
    return m.data.astype(np.complex128), m.indices, m.indptr


def _tol(dtype):
    return 1e-12 if dtype == np.complex128 else 1e-5


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("avg", [2, 12, 50, 150, 300])
@pytest.mark.parametrize("idx", [np.int32, np.int64])
def test_csr_expval_lane_widths(capi, dtype, avg, idx):
    rng = np.random.default_rng(1000 + avg)
    n = 10
    d, i, p = _random_csr(n, avg, rng)
    i, p = i.astype(idx), p.astype(idx)
    psi = random_state(n, rng)
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    ref = O.expval_csr(psi, d, i.astype(np.int64), p.astype(np.int64))
    scale = max(1.0, np.abs(d).sum() / (1 << n))
    assert abs(s.expval_csr(d, i, p) - ref) < _tol(dtype) * scale * 10
    ob = capi.Obs.sparse(d, i, p, list(range(n)))
    assert abs(s.expval(ob) - ref) < _tol(dtype) * scale * 10
    # Second use: the device copy cached on the observable.
    assert abs(s.expval(ob) - ref) < _tol(dtype) * scale * 10


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("avg", [3, 40, 150])
@pytest.mark.parametrize("idx", [np.int32, np.int64])
def test_csr_apply(capi, dtype, avg, idx):
    rng = np.random.default_rng(2000 + avg)
    n = 9
    d, i, p = _random_csr(n, avg, rng)
    psi = random_state(n, rng)
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    s.apply_obs(capi.Obs.sparse(d, i.astype(idx), p.astype(idx), list(range(n))))
    ref = O.spmv(psi, d, i.astype(np.int64), p.astype(np.int64))
    np.testing.assert_allclose(s.download(), ref, atol=_tol(dtype) * max(1.0, avg) * 10)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("idx", [np.int32, np.int64])
def test_adjoint_with_sparse_observable(capi, dtype, idx):
    rng = np.random.default_rng(3000)
    n = 8
    d, i, p = _random_csr(n, 6, rng, empty_rows=False)
    # Hermitian: A + A^dagger.
    a = sp.csr_matrix((d, i, p), shape=(1 << n, 1 << n))
    h = (a + a.conj().T).tocsr()
    h.sort_indices()
    hd, hi, hp = h.data.astype(np.complex128), h.indices.astype(idx), h.indptr.astype(idx)
    ops = random_ops(n, 40, rng, names=["RX", "RY", "RZ", "CNOT", "IsingXX", "CRY"])
    npar = sum(1 for q in ops[1] if q)
    tp = list(range(npar))
    psi = random_state(n, rng)
    ref = O.adjoint_jacobian(psi, [O.Sparse(hd, hi.astype(np.int64), hp.astype(np.int64), list(range(n))),
                                   O.Named("PauliZ", [2])], O.Ops(*ops), tp, True)
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    obs = [capi.Obs.sparse(hd, hi, hp, list(range(n))), capi.Obs.named("PauliZ", [2])]
    got = capi.adjoint_jacobian(s, obs, capi.Ops(*ops), tp, True)
    tol = 1e-10 if dtype == np.complex128 else 5e-5 * max(1.0, np.abs(hd).max())
    np.testing.assert_allclose(got, ref, atol=tol)


def test_csr_large_rows_and_int32_pauli(capi):
    """A 16-qubit Pauli-sum CSR (49-64 entries a row, 64-lane rows) in int32."""
    from helpers import pauli_csr, random_hamiltonian, to_oracle

    rng = np.random.default_rng(3100)
    n = 16
    ham = random_hamiltonian(n, 64, rng)
    d, i, p = pauli_csr(n, ham, index_dtype=np.int32)
    psi = random_state(n, rng)
    ref = O.expval(psi, to_oracle(ham))
    for dtype in (np.complex128, np.complex64):
        s = capi.State(n, dtype)
        s.upload(psi.astype(dtype))
        tol = 1e-10 * max(1.0, abs(ref)) if dtype == np.complex128 else 1e-4
        assert abs(s.expval_csr(d, i, p) - ref) < tol
        assert abs(s.expval(capi.Obs.sparse(d, i, p, list(range(n)))) - ref) < tol


@pytest.mark.parametrize("idx", [np.int32, np.int64])
def test_csr_malformed_is_reported(capi, idx):
    """Malformed matrices are reported by the kernels' own checks (nothing is
    read outside the arrays) and leave the state usable."""
    rng = np.random.default_rng(3200)
    n = 6
    N = 1 << n
    d, i, p = _random_csr(n, 4, rng, empty_rows=False)
    psi = random_state(n, rng)
    s = capi.State(n)
    s.upload(psi)
    bad_col = i.astype(idx).copy()
    bad_col[len(bad_col) // 2] = N + 5
    with pytest.raises(capi.PLException, match="outside the matrix"):
        s.expval_csr(d, bad_col, p.astype(idx))
    neg_col = i.astype(idx).copy()
    neg_col[3] = -1
    with pytest.raises(capi.PLException, match="outside the matrix"):
        s.expval_csr(d, neg_col, p.astype(idx))
    bad_ptr = p.astype(idx).copy()
    bad_ptr[10], bad_ptr[11] = bad_ptr[11] + 1, bad_ptr[10]  # decreasing range
    with pytest.raises(capi.PLException, match="outside the matrix"):
        s.expval_csr(d, i.astype(idx), bad_ptr)
    with pytest.raises(capi.PLException, match="outside the matrix"):
        s.apply_obs(capi.Obs.sparse(d, bad_col, p.astype(idx), list(range(n))))
    short = p.astype(idx).copy()
    short[-1] -= 1
    with pytest.raises(capi.PLException, match="number of non-zeros"):
        s.expval_csr(d, i.astype(idx), short)
    s.upload(psi)
    ref = O.expval_csr(psi, d, i.astype(np.int64), p.astype(np.int64))
    assert abs(s.expval_csr(d, i.astype(idx), p.astype(idx)) - ref) < 1e-12
