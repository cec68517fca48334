"""Transcribes golden vectors from the reference's C++ unit tests into JSON.

Reads, as text, the Catch2 tests of PennyLane-Lightning-Kokkos under
/root/reference/pennylane_lightning_kokkos/src/tests and writes
reference_goldens.json next to this script: input states, operation,
expected output vectors and expected scalars, each with the test file and
line range it came from. Only numbers are copied (fixtures), no code.

Run in the container that has /root/reference: python transcribe_reference_goldens.py
"""
import json
import os
import re

REF = "/root/reference/pennylane_lightning_kokkos/src/tests"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_goldens.json")

NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
PAIR = re.compile(r"\{\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*\}")


def lines(fname):
    with open(os.path.join(REF, fname)) as f:
        return f.read().split("\n")


def block(fname, start_pat, after=0):
    """Text of the brace-balanced initializer that follows `start_pat`."""
    L = lines(fname)
    text = "\n".join(L)
    idx = text.index(start_pat, after)
    j = text.index("{", idx + len(start_pat) - 1)
    depth = 0
    for k in range(j, len(text)):
        if text[k] == "{":
            depth += 1
        elif text[k] == "}":
            depth -= 1
            if depth == 0:
                line0 = text.count("\n", 0, idx) + 1
                line1 = text.count("\n", 0, k) + 1
                return text[j:k + 1], (line0, line1), k
    raise ValueError(start_pat)


def pairs(s):
    return [[float(a), float(b)] for a, b in PAIR.findall(s)]


def case_offset(fname, title):
    text = "\n".join(lines(fname))
    return text.index(title)


def gate_case_ini_expected(fname, title, gate, wires, params, n):
    off = case_offset(fname, title)
    ini, r0, _ = block(fname, "std::vector<cp_t> ini_st{", off)
    exp, r1, _ = block(fname, "std::vector<cp_t> expected{", off)
    return {"kind": "gate", "gate": gate, "wires": wires, "params": params, "inverse": False,
            "num_qubits": n, "init": pairs(ini), "expected": pairs(exp),
            "source": f"src/tests/{fname}:{r0[0]}-{r1[1]}"}


def sparse_assign(fname, title, name):
    """expected_results[i][k] = cp_t{x, y}; statements inside one test case."""
    off = case_offset(fname, title)
    text = "\n".join(lines(fname))
    end = text.find("TEMPLATE_TEST_CASE", off + 10)
    seg = text[off:end if end > 0 else None]
    pat = re.compile(re.escape(name) + r"\[(\d+)\]\[(\d+)\]\s*=\s*cp_t\{\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*\}")
    out = {}
    for m in pat.finditer(seg):
        out.setdefault(int(m.group(1)), {})[int(m.group(2))] = [float(m.group(3)), float(m.group(4))]
    line0 = text.count("\n", 0, off) + 1
    line1 = text.count("\n", 0, off + len(seg)) + 1
    return out, (line0, line1)


def dense(entries, dim):
    v = [[0.0, 0.0] for _ in range(dim)]
    for k, val in entries.items():
        v[k] = val
    return v


def main():
    goldens = []
    P = "Test_StateVectorKokkos_Param.cpp"
    # ini_st / expected pairs (IsingXY, excitations).
    goldens.append(gate_case_ini_expected(P, "applyIsingXY", "IsingXY", [0, 1], [0.312], 4))
    goldens.append(gate_case_ini_expected(P, "applySingleExcitation\"", "SingleExcitation", [0, 2],
                                          [0.267030328057308], 3))
    goldens.append(gate_case_ini_expected(P, "applySingleExcitationMinus", "SingleExcitationMinus",
                                          [0, 2], [0.267030328057308], 3))
    goldens.append(gate_case_ini_expected(P, "applySingleExcitationPlus", "SingleExcitationPlus",
                                          [0, 2], [0.267030328057308], 3))
    for g in ("DoubleExcitation\"", "DoubleExcitationMinus", "DoubleExcitationPlus"):
        goldens.append(gate_case_ini_expected(P, "apply" + g, g.strip('"'), [0, 1, 2, 3],
                                              [0.267030328057308], 4))
    # RX on |0> (expected_results / _adj for angles 0.1, 0.6).
    off = case_offset(P, "applyRX")
    exp, r, _ = block(P, "expected_results{", off)
    adj, r2, _ = block(P, "expected_results_adj{", off)
    e, a = pairs(exp), pairs(adj)
    for i, ang in enumerate([0.1, 0.6]):
        for inv, src in ((False, e), (True, a)):
            goldens.append({"kind": "gate", "gate": "RX", "wires": [0], "params": [ang], "inverse": inv,
                            "num_qubits": 1, "init": [[1, 0], [0, 0]], "expected": src[2 * i:2 * i + 2],
                            "source": f"src/tests/{P}:{r[0]}-{r2[1]}"})
    # RY on ini_st for angles 0.2, 0.7, 2.9.
    off = case_offset(P, "applyRY")
    exp, r, _ = block(P, "expected_results{", off)
    adj, _, _ = block(P, "expected_results_adj{", off)
    ini, r3, _ = block(P, "std::vector<cp_t> ini_st{", off)
    e, a, i0 = pairs(exp), pairs(adj), pairs(ini)
    for i, ang in enumerate([0.2, 0.7, 2.9]):
        for inv, src in ((False, e), (True, a)):
            goldens.append({"kind": "gate", "gate": "RY", "wires": [0], "params": [ang], "inverse": inv,
                            "num_qubits": 1, "init": i0, "expected": src[2 * i:2 * i + 2],
                            "source": f"src/tests/{P}:{r[0]}-{r3[1]}"})
    # Ising XX / YY / ZZ and MultiRZ on |000>: sparse expected assignments.
    for gate, pairs_w in (("IsingXX", [[0, 1], [0, 2]]), ("IsingYY", [[0, 1], [0, 2]]),
                          ("IsingZZ", [[0, 1], [0, 2]]), ("MultiRZ", [[0, 1], [0, 2]])):
        fwd, rr = sparse_assign(P, "apply" + gate, "expected_results")
        bwd, _ = sparse_assign(P, "apply" + gate, "expected_results_adj")
        angles = [0.3, 0.8]
        for inv, table in ((False, fwd), (True, bwd)):
            for wi, w in enumerate(pairs_w):
                for ai, ang in enumerate(angles):
                    idx = ai + (len(angles) * wi if len(table) > 2 else 0)
                    if idx not in table:
                        continue
                    goldens.append({"kind": "gate", "gate": gate, "wires": w, "params": [ang],
                                    "inverse": inv, "num_qubits": 3, "init": dense({0: [1, 0]}, 8),
                                    "expected": dense(table[idx], 8),
                                    "source": f"src/tests/{P}:{rr[0]}-{rr[1]}"})

    # Expectation values (Test_StateVectorKokkos_Expval.cpp).
    E = "Test_StateVectorKokkos_Expval.cpp"
    off = case_offset(E, "Hamiltonian_expval\"")
    init, r0, k = block(E, "std::vector<cp_t> init_state{", off)
    mat, r1, _ = block(E, "std::vector<cp_t> matrix{", k)
    goldens.append({"kind": "expval_matrix", "num_qubits": 3, "init": pairs(init), "wires": [0, 1, 2],
                    "matrix": pairs(mat), "expected": 1.263, "rel": 1e-7,
                    "source": f"src/tests/{E}:{r0[0]}-{r1[1]}"})
    off = case_offset(E, "Hamiltonian_expval_Sparse")
    init, r0, k = block(E, "std::vector<cp_t> init_state{", off)
    ptr, _, k = block(E, "index_ptr = {", k)
    ind, _, k = block(E, "indices = {", k)
    val, r1, _ = block(E, "values = {", k)
    goldens.append({"kind": "expval_csr", "num_qubits": 3, "init": pairs(init),
                    "indptr": [int(float(x)) for x in re.findall(NUM, ptr)],
                    "indices": [int(float(x)) for x in re.findall(NUM, ind)],
                    "data": pairs(val), "expected": 3.1415, "rel": 1e-7,
                    "source": f"src/tests/{E}:{r0[0]}-{r1[1]}"})
    for title, ob, expected in (("Test expectation value of HamiltonianObs",
                                 ["ham", [0.3, 0.5], [["named", "PauliX", [0]], ["named", "PauliZ", [1]]]],
                                 -0.086),
                                ("Test expectation value of TensorProdObs",
                                 ["tensor", [["named", "PauliX", [0]], ["named", "PauliZ", [1]]]], -0.36)):
        off = case_offset(E, title)
        init, r0, _ = block(E, "init_state{", off)
        goldens.append({"kind": "expval_obs", "num_qubits": 3, "init": pairs(init), "obs": ob,
                        "expected": expected, "rel": 1.2e-5, "source": f"src/tests/{E}:{r0[0]}-{r0[1] + 14}"})

    # Linear algebra (Test_LinearAlgebra.cpp).
    A = "Test_LinearAlgebra.cpp"
    off = case_offset(A, "Linear Algebra::SparseMV")
    vec, r0, k = block(A, "vectors = {", off)
    ref, _, k = block(A, "result_refs = {", k)
    ptr, _, k = block(A, "indptr = {", k)
    ind, _, k = block(A, "indices = {", k)
    val, r1, _ = block(A, "values = {", k)
    goldens.append({"kind": "spmv", "x": pairs(vec), "expected": pairs(ref),
                    "indptr": [int(float(x)) for x in re.findall(NUM, ptr)],
                    "indices": [int(float(x)) for x in re.findall(NUM, ind)],
                    "data": pairs(val), "source": f"src/tests/{A}:{r0[0]}-{r1[1]}"})
    off = case_offset(A, "Linear Algebra::axpy_Kokkos")
    alpha, r0, k = block(A, "cp_t alpha = {", off)
    v0, _, k = block(A, "v0 = {", k)
    v1, _, k = block(A, "v1 = {", k)
    ref, r1, _ = block(A, "result_refs = {", k)
    goldens.append({"kind": "axpy", "alpha": [float(x) for x in re.findall(NUM, alpha)],
                    "x": pairs(v0), "y": pairs(v1), "expected": pairs(ref),
                    "source": f"src/tests/{A}:{r0[0]}-{r1[1]}"})

    with open(OUT, "w") as f:
        json.dump({"reference": "PennyLane-Lightning-Kokkos 0.33.0-dev", "goldens": goldens}, f, indent=1)
    print(f"wrote {len(goldens)} goldens to {OUT}")


if __name__ == "__main__":
    main()
