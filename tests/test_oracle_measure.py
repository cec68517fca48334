"""Pins of the oracle's probabilities and sample selection.

probs: the golden tables of src/tests/Test_StateVectorKokkos_Measure.cpp:21-46
(every wire order on the reference's Initializing_StateVector) and of
tests/test_probs.py, plus numpy marginalisation of |psi|^2 on random states
and wire orders. Sample selection: for given uniforms the chosen basis index
k must bracket u between the exclusive CDF values (cdf[k] < u <= cdf[k+1]),
and a basis state is always sampled. The uniform generator (SplitMix64) is
pinned by its published first output for seed 0."""
import numpy as np
import pytest

import oracle as O
from helpers import initializing_state, random_state, sample_uniforms, splitmix64

MEASURE_GOLDENS = [
    ([0, 1, 2], [0.67078706, 0.03062806, 0.0870997, 0.00397696, 0.17564072, 0.00801973,
                 0.02280642, 0.00104134]),
    ([0, 2, 1], [0.67078706, 0.0870997, 0.03062806, 0.00397696, 0.17564072, 0.02280642,
                 0.00801973, 0.00104134]),
    ([1, 0, 2], [0.67078706, 0.03062806, 0.17564072, 0.00801973, 0.0870997, 0.00397696,
                 0.02280642, 0.00104134]),
    ([1, 2, 0], [0.67078706, 0.0870997, 0.17564072, 0.02280642, 0.03062806, 0.00397696,
                 0.00801973, 0.00104134]),
    ([2, 0, 1], [0.67078706, 0.17564072, 0.03062806, 0.00801973, 0.0870997, 0.02280642,
                 0.00397696, 0.00104134]),
    ([2, 1, 0], [0.67078706, 0.17564072, 0.0870997, 0.02280642, 0.03062806, 0.00801973,
                 0.00397696, 0.00104134]),
    ([0, 1], [0.70141512, 0.09107666, 0.18366045, 0.02384776]),
    ([0, 2], [0.75788676, 0.03460502, 0.19844714, 0.00906107]),
    ([1, 2], [0.84642778, 0.0386478, 0.10990612, 0.0050183]),
    ([2, 1], [0.84642778, 0.10990612, 0.0386478, 0.0050183]),
    ([0], [0.79249179, 0.20750821]),
    ([1], [0.88507558, 0.11492442]),
    ([2], [0.9563339, 0.0436661]),
]


@pytest.mark.parametrize("wires, expected", MEASURE_GOLDENS)
def test_probs_reference_goldens(wires, expected):
    sv = initializing_state(3)
    np.testing.assert_allclose(O.probs(sv, wires), expected, atol=1e-6)
    np.testing.assert_allclose(numpy_probs(sv, reference_order(wires)), expected, atol=1e-6)


def _py_probs_circuits():
    sv1 = O.zero_state(2)
    O.apply(sv1, "RX", [0], False, [0.4])
    for g, a in (("RZ", 0.5), ("RY", 0.3), ("RZ", -0.7)):  # Rot(0.5, 0.3, -0.7)
        O.apply(sv1, g, [0], False, [a])
    O.apply(sv1, "RY", [0], False, [-0.2])
    sv2 = O.zero_state(2)
    O.apply(sv2, "RX", [0], False, [0.5])
    O.apply(sv2, "RY", [1], False, [0.3])
    return sv1, sv2


def test_probs_python_goldens():
    """tests/test_probs.py:87-191 (the [1, 0] orders are the values that test
    lists for out-of-order wires)."""
    sv1, sv2 = _py_probs_circuits()
    np.testing.assert_allclose(O.probs(sv1, [0, 1]), [0.9165164490394898, 0.0, 0.08348355096051052, 0.0],
                               atol=1e-12)
    np.testing.assert_allclose(O.probs(sv1, [0]), [0.9165164490394898, 0.08348355096051052], atol=1e-12)
    np.testing.assert_allclose(O.probs(sv1, [1, 0]), [0.9165164490394898, 0.08348355096051052, 0.0, 0.0],
                               atol=1e-12)
    np.testing.assert_allclose(O.probs(sv2, [0, 1]), [0.9178264236525453, 0.02096485729264079,
                                                      0.059841820910257436, 0.0013668981445561978],
                               atol=1e-12)
    np.testing.assert_allclose(O.probs(sv2, [0]), [0.938791280945186, 0.061208719054813635], atol=1e-12)
    np.testing.assert_allclose(O.probs(sv2, [1, 0]), [0.9178264236525453, 0.059841820910257436,
                                                      0.02096485729264079, 0.0013668981445561978],
                               atol=1e-12)


def reference_order(wires):
    """The reference's transposition (MeasuresKokkos.hpp:487-512,
    getTransposedIndexFunctor) returns the plain marginal over the wire order
    wires[pi[pi[k]]], pi = argsort(wires): the given order itself for
    involutive orders, its inverse-permuted form for cyclic ones (the golden
    rows [1, 2, 0] and [2, 0, 1] above fix this)."""
    pi = sorted(range(len(wires)), key=lambda k: wires[k])
    return [wires[pi[pi[k]]] for k in range(len(wires))]


def numpy_probs(sv, wires):
    n = int(np.log2(sv.size))
    p = (np.abs(sv) ** 2).reshape((2,) * n)
    rest = tuple(q for q in range(n) if q not in wires)
    m = p.sum(axis=rest) if rest else p
    kept = [q for q in range(n) if q in wires]
    return np.transpose(m, [kept.index(w) for w in wires]).reshape(-1)


@pytest.mark.parametrize("seed", range(8))
def test_probs_numpy_marginals(seed):
    rng = np.random.default_rng(500 + seed)
    n = 2 + seed % 6
    sv = random_state(n, rng)
    k = int(rng.integers(1, n + 1))
    wires = [int(x) for x in rng.permutation(n)[:k]]
    np.testing.assert_allclose(O.probs(sv, wires), numpy_probs(sv, reference_order(wires)),
                               atol=1e-14)
    np.testing.assert_allclose(O.probs(sv, sorted(wires)), numpy_probs(sv, sorted(wires)),
                               atol=1e-14)
    np.testing.assert_allclose(O.probs(sv), np.abs(sv) ** 2, atol=1e-15)
    assert abs(O.probs(sv, wires).sum() - 1.0) < 1e-12


def test_sample_indices_bracket_uniforms():
    rng = np.random.default_rng(7)
    sv = random_state(6, rng)
    u = rng.uniform(size=5000)
    u[0] = 1.0
    idx = O.sample_indices(sv, u)
    cdf = np.concatenate([[0.0], np.cumsum(np.abs(sv) ** 2)])
    cdf[-1] = max(cdf[-1], 1.0)
    for k, x in zip(idx, u):
        assert (cdf[k] < x or k == 0) and x <= cdf[k + 1] + 1e-15


def test_sample_indices_basis_state():
    sv = np.zeros(32, dtype=np.complex128)
    sv[19] = 1.0
    assert set(O.sample_indices(sv, np.linspace(1e-9, 1.0, 257))) == {19}


def test_splitmix64_known_answer():
    """First output of SplitMix64 from state 0 (the published reference
    implementation's test value)."""
    assert splitmix64(1, 0) == 0xE220A8397B1DCDAF
    u = sample_uniforms(12345, 1000)
    assert np.all(u > 0) and np.all(u <= 1) and abs(u.mean() - 0.5) < 0.05
