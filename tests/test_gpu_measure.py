"""probs and sampling on the GPU against the oracle (MeasuresKokkos.hpp:389-569).

probs: FP64 marginals for sorted, unsorted, single and full wire lists, both
precisions (the oracle reads the state the GPU holds, so complex64 compares
at FP64 summation error), the reference's golden table through the pybind
module. Samples: the basis indices equal the oracle's selection for the same
SplitMix64 uniforms; a basis state is always sampled; a chi-square test of
200 000 shots against the probabilities; the pybind GenerateSamples layout."""
import numpy as np
import pytest

import oracle as O
from helpers import initializing_state, random_ops, random_state, sample_uniforms
from test_oracle_measure import MEASURE_GOLDENS

pytestmark = pytest.mark.gpu


def _state(capi, n, dtype, seed):
    rng = np.random.default_rng(seed)
    psi = random_state(n, rng)
    s = capi.State(n, dtype)
    s.upload(psi.astype(dtype))
    return s, s.download().astype(np.complex128), rng


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [3, 9, 13, 17])
def test_probs_match_oracle(capi, dtype, n):
    s, psi, rng = _state(capi, n, dtype, 40 + n)
    cases = [[], list(range(n)), [n - 1], [0], sorted(int(x) for x in rng.permutation(n)[: n // 2]),
             [int(x) for x in rng.permutation(n)[: max(1, n - 1)]],
             [int(x) for x in rng.permutation(n)]]
    for wires in cases:
        got = s.probs(wires)
        ref = O.probs(psi, wires)
        np.testing.assert_allclose(got, ref, atol=1e-12, err_msg=f"n={n} wires={wires}")


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_probs_reference_goldens_module(dtype):
    from pennylane_lightning_kokkos_amd import lightning_kokkos_qubit_ops as m

    sv = (m.LightningKokkos_C128 if dtype == np.complex128 else m.LightningKokkos_C64)(3)
    a = 0.7
    for w in range(3):
        sv.RX([w], False, [a])
        sv.RY([w], False, [a])
        a -= 0.2
    for wires, expected in MEASURE_GOLDENS:
        got = sv.probs(wires)
        assert got.dtype == (np.float64 if dtype == np.complex128 else np.float32)
        np.testing.assert_allclose(got, expected, atol=1e-6)


def test_probs_after_circuit_24q_marginal(capi):
    n = 22
    rng = np.random.default_rng(3)
    ops = random_ops(n, 60, rng, names=["RX", "RY", "CNOT", "IsingXX", "Hadamard"])
    s = capi.State(n)
    s.apply_ops(capi.Ops(*ops))
    psi = s.download()
    for wires in ([5, 17, 2], [21], [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10]):
        np.testing.assert_allclose(s.probs(wires), O.probs(psi, wires), atol=1e-12)
    assert abs(s.probs([3]).sum() - 1.0) < 1e-12


def _indices(rows):
    n = rows.shape[1]
    return (rows.astype(np.int64) << np.arange(n - 1, -1, -1)).sum(axis=1)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_samples_match_oracle_selection(capi, dtype):
    n, shots, seed = 10, 5000, 99
    s, psi, _ = _state(capi, n, dtype, 77)
    rows = s.generate_samples(shots, seed)
    assert rows.shape == (shots, n) and set(np.unique(rows)) <= {0, 1}
    ref = O.sample_indices(psi, sample_uniforms(seed, shots))
    np.testing.assert_array_equal(_indices(rows), ref)


def test_samples_basis_state(capi):
    s = capi.State(8)
    s.set_basis(37)
    rows = s.generate_samples(2000, 5)
    assert set(_indices(rows)) == {37}


def test_samples_chi_square(capi):
    from scipy.stats import chi2

    n, shots = 4, 200_000
    s, psi, _ = _state(capi, n, np.complex128, 11)
    counts = np.bincount(_indices(s.generate_samples(shots, 2024)), minlength=2**n)
    expected = shots * np.abs(psi) ** 2
    stat = ((counts - expected) ** 2 / expected).sum()
    assert chi2.sf(stat, 2**n - 1) > 1e-6, (stat, counts, expected)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_generate_samples_module(dtype):
    from pennylane_lightning_kokkos_amd import lightning_kokkos_qubit_ops as m

    sv = (m.LightningKokkos_C128 if dtype == np.complex128 else m.LightningKokkos_C64)(2)
    sv.RX([0], False, [1.5708])
    smp = np.array(sv.GenerateSamples(2, 1000))
    assert smp.shape == (1000, 2) and set(np.unique(smp)) <= {0, 1}
    assert not smp[:, 1].any() and 350 < smp[:, 0].sum() < 650


def test_probs_and_samples_argument_errors(capi):
    s = capi.State(4)
    with pytest.raises(capi.PLException, match="repeated wire"):
        s.probs([1, 1])
    with pytest.raises(capi.PLException, match="out of range"):
        s.probs([4])
    assert s.generate_samples(0).shape == (0, 4)
    np.testing.assert_allclose(s.probs([3, 0]), [1, 0, 0, 0])
