"""Circuit, state and observable builders shared by the CPU and GPU tests.

Observables are described by plain tuples so the same description can be
turned into an oracle object (checker) and into a C-ABI handle (product):
  ("named", name, wires) | ("herm", matrix, wires) | ("tensor", [obs...])
  | ("ham", coeffs, [obs...]) | ("sparse", data, indices, indptr, wires)
"""
from __future__ import annotations

import numpy as np

import oracle as O

# name -> (wires, params); wires 0 means "any number" (MultiRZ)
GATES = {
    "Identity": (1, 0), "PauliX": (1, 0), "PauliY": (1, 0), "PauliZ": (1, 0),
    "Hadamard": (1, 0), "S": (1, 0), "T": (1, 0), "PhaseShift": (1, 1),
    "RX": (1, 1), "RY": (1, 1), "RZ": (1, 1), "Rot": (1, 3),
    "CNOT": (2, 0), "SWAP": (2, 0), "CY": (2, 0), "CZ": (2, 0),
    "ControlledPhaseShift": (2, 1), "CRX": (2, 1), "CRY": (2, 1), "CRZ": (2, 1),
    "CRot": (2, 3), "IsingXX": (2, 1), "IsingXY": (2, 1), "IsingYY": (2, 1),
    "IsingZZ": (2, 1), "SingleExcitation": (2, 1), "SingleExcitationMinus": (2, 1),
    "SingleExcitationPlus": (2, 1), "CSWAP": (3, 0), "Toffoli": (3, 0),
    "DoubleExcitation": (4, 1), "DoubleExcitationMinus": (4, 1),
    "DoubleExcitationPlus": (4, 1), "MultiRZ": (0, 1),
}

# Parametric gates with a generator (one parameter each).
GEN_GATES = [g for g, (w, p) in GATES.items() if p == 1]


def random_state(n: int, rng: np.random.Generator) -> np.ndarray:
    v = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    return (v / np.linalg.norm(v)).astype(np.complex128)


def random_unitary(dim: int, rng: np.random.Generator) -> np.ndarray:
    z = (rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))


def random_hermitian(dim: int, rng: np.random.Generator) -> np.ndarray:
    a = rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))
    return (a + a.conj().T) / 2


def gate_wires(name: str, n: int, rng: np.random.Generator, k_multi: int = 3):
    w, _ = GATES[name]
    if w == 0:
        w = min(n, k_multi)
    return [int(x) for x in rng.choice(n, size=w, replace=False)]


def random_ops(n: int, count: int, rng: np.random.Generator, names=None,
               allow_inverse=True, with_matrix=False):
    """Random op list (single-parameter gates only, as the adjoint requires)."""
    names = names or [g for g in GATES if GATES[g][1] <= 1 and GATES[g][0] <= n]
    ops_n, ops_p, ops_w, ops_i, ops_m = [], [], [], [], []
    for _ in range(count):
        g = names[rng.integers(len(names))]
        w = gate_wires(g, n, rng)
        p = [float(rng.uniform(0, 2 * np.pi))] if GATES[g][1] == 1 else []
        ops_n.append(g)
        ops_p.append(p)
        ops_w.append(w)
        ops_i.append(bool(allow_inverse and rng.integers(2)))
        ops_m.append([])
    if with_matrix:
        k = int(rng.integers(1, min(3, n) + 1))
        w = [int(x) for x in rng.choice(n, size=k, replace=False)]
        pos = int(rng.integers(len(ops_n) + 1))
        ops_n.insert(pos, "QubitUnitary")
        ops_p.insert(pos, [])
        ops_w.insert(pos, w)
        ops_i.insert(pos, bool(allow_inverse and rng.integers(2)))
        ops_m.insert(pos, random_unitary(2**k, rng).reshape(-1))
    return ops_n, ops_p, ops_w, ops_i, ops_m


def random_pauli_word(n: int, rng, max_weight=4):
    k = int(rng.integers(1, min(max_weight, n) + 1))
    wires = sorted(int(x) for x in rng.choice(n, size=k, replace=False))
    names = [("PauliX", "PauliY", "PauliZ")[rng.integers(3)] for _ in wires]
    if k == 1:
        return ("named", names[0], [wires[0]])
    return ("tensor", [("named", nm, [w]) for nm, w in zip(names, wires)])


def random_hamiltonian(n: int, terms: int, rng, max_weight=4):
    coeffs = [float(c) for c in rng.normal(size=terms)]
    return ("ham", coeffs, [random_pauli_word(n, rng, max_weight) for _ in range(terms)])


def to_oracle(ob):
    kind = ob[0]
    if kind == "named":
        return O.Named(ob[1], list(ob[2]))
    if kind == "herm":
        return O.Hermitian(np.asarray(ob[1]), list(ob[2]))
    if kind == "tensor":
        return O.Tensor([to_oracle(f) for f in ob[1]])
    if kind == "ham":
        return O.Hamiltonian(list(ob[1]), [to_oracle(t) for t in ob[2]])
    if kind == "sparse":
        return O.Sparse(ob[1], ob[2], ob[3], list(ob[4]))
    raise ValueError(kind)


def to_capi(capi, ob):
    kind = ob[0]
    if kind == "named":
        return capi.Obs.named(ob[1], ob[2])
    if kind == "herm":
        return capi.Obs.hermitian(ob[1], ob[2])
    if kind == "tensor":
        return capi.Obs.tensor([to_capi(capi, f) for f in ob[1]])
    if kind == "ham":
        return capi.Obs.hamiltonian(ob[1], [to_capi(capi, t) for t in ob[2]])
    if kind == "sparse":
        return capi.Obs.sparse(ob[1], ob[2], ob[3], ob[4])
    raise ValueError(kind)


def dense_pauli_sparse(n: int, ham, rng=None):
    """CSR (data, indices, indptr) of a Hamiltonian of Pauli words (numpy)."""
    import scipy.sparse as sp

    def word_matrix(ob):
        mats = {"PauliX": np.array([[0, 1], [1, 0]], complex),
                "PauliY": np.array([[0, -1j], [1j, 0]], complex),
                "PauliZ": np.array([[1, 0], [0, -1]], complex),
                "Identity": np.eye(2, dtype=complex)}
        factors = {}
        if ob[0] == "named":
            factors[ob[2][0]] = mats[ob[1]]
        else:
            for f in ob[1]:
                factors[f[2][0]] = mats[f[1]]
        M = sp.identity(1, dtype=complex, format="csr")
        for w in range(n):
            M = sp.kron(M, sp.csr_matrix(factors.get(w, np.eye(2, dtype=complex))), format="csr")
        return M

    H = None
    for c, t in zip(ham[1], ham[2]):
        term = c * word_matrix(t)
        H = term if H is None else H + term
    H = sp.csr_matrix(H)
    H.sort_indices()
    return H.data.astype(np.complex128), H.indices.astype(np.int64), H.indptr.astype(np.int64)


# ----------------------------------------------------------------------------
# Benchmark workloads (seeded).
# ----------------------------------------------------------------------------
def c1_workload(seed=20261202):
    """12 qubits; 4 layers of RX, RY, RZ on every wire, then CNOT ring; Z on each wire."""
    rng = np.random.default_rng(seed)
    n = 12
    names, params, wires = [], [], []
    for _ in range(4):
        for g in ("RX", "RY", "RZ"):
            for w in range(n):
                names.append(g)
                params.append([float(rng.uniform(0, 2 * np.pi))])
                wires.append([w])
        for w in range(n):
            names.append("CNOT")
            params.append([])
            wires.append([w, (w + 1) % n])
    obs = [("named", "PauliZ", [w]) for w in range(n)]
    return n, (names, params, wires, [False] * len(names), [[] for _ in names]), obs


def c2_workload(n=26, count=2000, seed=20261203, ham_terms=64):
    """Gate-throughput stream: `count` gates drawn uniformly from
    C2_GATES on uniform distinct wires, angles uniform in [0, 2 pi); then a
    Hamiltonian of `ham_terms` random Pauli words of weight 1-4 with N(0, 1)
    coefficients."""
    rng = np.random.default_rng(seed)
    ops = random_ops(n, count, rng, names=C2_GATES, allow_inverse=False)
    ham = random_hamiltonian(n, ham_terms, rng)
    return n, ops, ham


C2_GATES = ["Hadamard", "RX", "RY", "RZ", "CNOT", "CZ", "CRX", "SWAP", "IsingXX", "IsingZZ"]


def pauli_words(n, ham):
    """(x mask, z mask, coefficient) per term of a Hamiltonian of Pauli words;
    wire w is bit n - 1 - w of the amplitude index."""
    out = []
    for c, t in zip(ham[1], ham[2]):
        factors = [t] if t[0] == "named" else t[1]
        x = z = 0
        ny = 0
        for f in factors:
            b = 1 << (n - 1 - f[2][0])
            if f[1] in ("PauliX", "PauliY"):
                x |= b
            if f[1] in ("PauliZ", "PauliY"):
                z |= b
            ny += f[1] == "PauliY"
        out.append((x, z, complex(c) * 1j ** ny))  # Y = i X Z on one wire
    return out


def pauli_csr(n, ham, index_dtype=np.int64):
    """CSR (data, indices, indptr) of a sum of Pauli words, built directly:
    the word c X^x Z^z (Y = i X Z per wire, folded into c) has, in row r,
    the single entry c (-1)^popcount(col & z) at column col = r ^ x. Rows
    hold one entry per distinct x mask, columns sorted. Numpy only, so
    20-qubit matrices build in seconds."""
    N = 1 << n
    rows = np.arange(N, dtype=np.int64)
    by_x = {}
    for x, z, c in pauli_words(n, ham):
        # (X^x Z^z)|col> = (-1)^popcount(col & z) |col ^ x>
        col = rows ^ x
        sign = 1.0 - 2.0 * (np.bitwise_count(col & z) & 1)
        by_x[x] = by_x.get(x, 0) + c * sign
    xs = sorted(by_x)
    cols = np.stack([rows ^ x for x in xs], axis=1)
    vals = np.stack([by_x[x] for x in xs], axis=1).astype(np.complex128)
    order = np.argsort(cols, axis=1, kind="stable")
    cols = np.take_along_axis(cols, order, axis=1)
    vals = np.take_along_axis(vals, order, axis=1)
    indptr = np.arange(0, N * len(xs) + 1, len(xs), dtype=index_dtype)
    return vals.reshape(-1), cols.reshape(-1).astype(index_dtype), indptr


def c3_workload(n=24, layers=10, seed=20261204, ham_terms=32):
    """Strongly entangling layers: Rot(phi, theta, omega) on every wire as
    RZ, RY, RZ, then CNOT(i, (i + r) mod n) with r = (layer mod (n-1)) + 1.
    Observables: PauliZ on wires 0..7 plus one random Pauli Hamiltonian."""
    rng = np.random.default_rng(seed)
    names, params, wires = [], [], []
    for layer in range(layers):
        for w in range(n):
            phi, theta, omega = (float(x) for x in rng.uniform(0, 2 * np.pi, size=3))
            for g, a in (("RZ", phi), ("RY", theta), ("RZ", omega)):
                names.append(g)
                params.append([a])
                wires.append([w])
        r = (layer % (n - 1)) + 1
        for w in range(n):
            names.append("CNOT")
            params.append([])
            wires.append([w, (w + r) % n])
    obs = [("named", "PauliZ", [w]) for w in range(min(8, n))]
    obs.append(random_hamiltonian(n, ham_terms, rng))
    return n, (names, params, wires, [False] * len(names), [[] for _ in names]), obs


def haar_su2(rng):
    """A Haar-random SU(2) matrix: QR of a complex Gaussian matrix with the
    phases of R's diagonal moved into Q (Mezzadri 2007), divided by a square
    root of its determinant."""
    z = (rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    q = q * (d / np.abs(d))
    return q / np.sqrt(np.linalg.det(q))


def c5_workload(n=32, layers=40, seed=20261207):
    """Large-state fused form: each layer applies a Haar SU(2) (as a matrix op
    named QubitUnitary) to every wire, then CZ on a brick pattern (pairs
    (0,1), (2,3), ... on even layers, (1,2), (3,4), ... on odd layers).
    Observable: the open-chain transverse-field Ising Hamiltonian
    sum_i Z_i Z_{i+1} + 0.5 sum_i X_i."""
    rng = np.random.default_rng(seed)
    names, params, wires, mats = [], [], [], []
    for layer in range(layers):
        for w in range(n):
            names.append("QubitUnitary")
            params.append([])
            wires.append([w])
            mats.append(haar_su2(rng).reshape(-1))
        for w in range(layer % 2, n - 1, 2):
            names.append("CZ")
            params.append([])
            wires.append([w, w + 1])
            mats.append([])
    coeffs = [1.0] * (n - 1) + [0.5] * n
    terms = ([("tensor", [("named", "PauliZ", [i]), ("named", "PauliZ", [i + 1])]) for i in range(n - 1)]
             + [("named", "PauliX", [i]) for i in range(n)])
    return n, (names, params, wires, [False] * len(names), mats), ("ham", coeffs, terms)


def excitation_workload(n=20, doubles=120, singles=60, n_words=32, seed=20261209):
    """A chemistry-style ansatz: the Hartree-Fock state (PauliX on the first
    n/2 wires, the occupied orbitals), then `doubles` DoubleExcitation gates
    on two occupied and two virtual wires and `singles` SingleExcitation gates
    on one of each, interleaved (two doubles, then one single), angles
    uniform in [0, 2 pi). Observables: one Hamiltonian of `n_words` random
    Pauli words and PauliZ on wires 0..3."""
    rng = np.random.default_rng(seed)
    occ = list(range(n // 2))
    virt = list(range(n // 2, n))
    names, params, wires = [], [], []
    for w in occ:
        names.append("PauliX")
        params.append([])
        wires.append([w])
    d = s_ = 0
    while d < doubles or s_ < singles:
        if (d + s_) % 3 == 2 and s_ < singles or d >= doubles:
            names.append("SingleExcitation")
            wires.append([int(rng.choice(occ)), int(rng.choice(virt))])
            s_ += 1
        else:
            names.append("DoubleExcitation")
            wires.append([int(x) for x in rng.choice(occ, 2, replace=False)] +
                         [int(x) for x in rng.choice(virt, 2, replace=False)])
            d += 1
        params.append([float(rng.uniform(0, 2 * np.pi))])
    ops = (names, params, wires, [False] * len(names), [[] for _ in names])
    obs = [random_hamiltonian(n, n_words, rng)] + [("named", "PauliZ", [w]) for w in range(4)]
    return n, ops, obs


def c4_workload(count=4096, seed=20261205, nmin=6, nmax=16, dmin=2, dmax=8):
    """Independent circuits: n uniform in [nmin, nmax], depth uniform in
    [dmin, dmax] layers of RY on every wire, RZ on every wire, then the CNOT
    chain (w, w + 1); parameters uniform in [0, 2 pi). Observable sum_i Z_i.
    Returns a list of (n, ops, obs)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        depth = int(rng.integers(dmin, dmax + 1))
        names, params, wires = [], [], []
        for _layer in range(depth):
            for g in ("RY", "RZ"):
                for w in range(n):
                    names.append(g)
                    params.append([float(rng.uniform(0, 2 * np.pi))])
                    wires.append([w])
            for w in range(n - 1):
                names.append("CNOT")
                params.append([])
                wires.append([w, w + 1])
        obs = ("ham", [1.0] * n, [("named", "PauliZ", [w]) for w in range(n)])
        out.append((n, (names, params, wires, [False] * len(names), [[] for _ in names]), obs))
    return out


# ----------------------------------------------------------------------------
# Sampling helpers.
# ----------------------------------------------------------------------------
SPLITMIX_GAMMA = 0x9E3779B97F4A7C15
MASK64 = (1 << 64) - 1


def splitmix64(k_plus_1: int, seed: int) -> int:
    """Output k of the SplitMix64 generator seeded with `seed` (Steele, Lea
    and Flood 2014): mix(seed + (k + 1) * gamma)."""
    z = (seed + k_plus_1 * SPLITMIX_GAMMA) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def sample_uniforms(seed: int, shots: int) -> np.ndarray:
    """The uniforms in (0, 1] that plk_generate_samples draws for each shot."""
    return np.array([((splitmix64(s + 1, seed) >> 11) + 1) * 2.0 ** -53 for s in range(shots)])


def initializing_state(n=3) -> np.ndarray:
    """Initializing_StateVector of the reference tests (TestHelpers.hpp:154-177):
    RX(a), RY(a) on each wire with a = 0.7, 0.5, 0.3, ..."""
    sv = O.zero_state(n)
    a = 0.7
    for w in range(n):
        O.apply(sv, "RX", [w], False, [a])
        O.apply(sv, "RY", [w], False, [a])
        a -= 0.2
    return sv
