"""Transcribes golden vectors from the reference's C++ unit tests into JSON.

Reads, as text, the Catch2 tests of PennyLane-Lightning-Kokkos under
/root/reference/pennylane_lightning_kokkos/src/tests and writes
reference_goldens.json next to this script: input states, operation,
expected output vectors and expected scalars, each with the test file and
line range it came from. Only numbers are copied (fixtures), no code.

Run in the container that has /root/reference: python transcribe_reference_goldens.py
"""
import ast
import cmath
import json
import math
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


# ---------------------------------------------------------------------------
# Symbolic expected tables: the NonParam tests and the RZ / PhaseShift /
# ControlledPhaseShift / Rot / CRot cases of the Param tests write their
# expected vectors with named constants (Util::INVSQRT2, ...) or with
# elements of host gate matrices (Gates::getRZ(a)[k], ...). The tables are
# read from the test text; constants and matrix elements are evaluated here
# from their mathematical definitions.
# ---------------------------------------------------------------------------
UTIL = {"HALF": 0.5, "INVSQRT2": 1 / math.sqrt(2), "IMAG": 1j, "ONE": 1.0, "ZERO": 0.0,
        "NEGONE": -1.0}
UTIL_CALL = re.compile(r"Util::(\w+)<[^>]*>\(\)")


def eval_expr(text, env):
    """Value of a restricted arithmetic expression (+, -, *, parentheses,
    numbers and names bound in env) taken from the test text."""
    tree = ast.parse(" ".join(UTIL_CALL.sub(r"\1", text).split()), mode="eval")

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult)):
            a, b = ev(node.left), ev(node.right)
            return a + b if isinstance(node.op, ast.Add) else a - b if isinstance(node.op, ast.Sub) else a * b
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
            return -ev(node.operand)
        if isinstance(node, ast.Name) and node.id in env:
            return env[node.id]
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return node.value
        raise ValueError(f"unsupported expression: {text!r}")

    return complex(ev(tree))


def case_text(fname, title):
    """(text, first line) of one TEMPLATE_TEST_CASE, up to the next one."""
    text = "\n".join(lines(fname))
    off = text.index(title)
    start = text.rfind("TEMPLATE_TEST_CASE", 0, off)
    end = text.find("TEMPLATE_TEST_CASE", off)
    return text[start:end if end > 0 else None], text.count("\n", 0, start) + 1


def bind_autos(text, env):
    """Evaluates every `auto name = expr;` of the test text into env, in order."""
    for m in re.finditer(r"(?:const\s+)?auto\s+(\w+)\s*=\s*([^;]+);", text):
        try:
            env[m.group(1)] = eval_expr(m.group(2), env)
        except (ValueError, SyntaxError):
            pass
    return env


def brace_block(text, start_pat, after=0):
    idx = text.index(start_pat, after)
    j = text.index("{", idx + len(start_pat) - 1)
    depth = 0
    for k in range(j, len(text)):
        depth += {"{": 1, "}": -1}.get(text[k], 0)
        if depth == 0:
            return text[j + 1:k], k
    raise ValueError(start_pat)


def split_top(s):
    """Top-level comma-separated items of an initializer body."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def cvec(values):
    return [[float(complex(v).real), float(complex(v).imag)] for v in values]


def rz_mat(a):
    """RZ(a) = diag(e^{-ia/2}, e^{ia/2}), row-major (GatesHost.hpp:405-410)."""
    return [cmath.exp(-0.5j * a), 0, 0, cmath.exp(0.5j * a)]


def phase_shift_mat(a):
    """PhaseShift(a) = diag(1, e^{ia}), row-major (GatesHost.hpp:312-317)."""
    return [1, 0, 0, cmath.exp(1j * a)]


def rot_mat(phi, theta, omega):
    """Rot = RZ(omega) RY(theta) RZ(phi) as written at GatesHost.hpp:431-435."""
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return [cmath.exp(-0.5j * (phi + omega)) * c, -cmath.exp(0.5j * (phi - omega)) * s,
            cmath.exp(-0.5j * (phi - omega)) * s, cmath.exp(0.5j * (phi + omega)) * c]


def numbers(s):
    return [float(x) for x in re.findall(NUM, s)]


H3 = [1 / (2 * math.sqrt(2))] * 8  # Hadamard on each of 3 wires from |000>


def param_closed_form_goldens():
    """RZ, PhaseShift, ControlledPhaseShift, Rot and CRot of
    Test_StateVectorKokkos_Param.cpp: tables of host-matrix elements."""
    P = "Test_StateVectorKokkos_Param.cpp"
    out = []
    elem = re.compile(r"(\w+)_data\[(\d+)\]\[(\d+)\]")
    for gate, mat_fn, title in (("RZ", rz_mat, "applyRZ\""), ("PhaseShift", phase_shift_mat, "applyPhaseShift\""),
                                ("ControlledPhaseShift", phase_shift_mat, "applyControlledPhaseShift\"")):
        text, l0 = case_text(P, title)
        ang_s, _ = brace_block(text, "angles{")
        angles = numbers(ang_s)
        body, _ = brace_block(text, "expected_results = {")
        rows = []
        for row in split_top(body):
            vals = []
            for item in split_top(row.strip()[1:-1]):
                m = elem.fullmatch(item.replace(" ", "").replace("\n", ""))
                vals.append(mat_fn(angles[int(m.group(2))])[int(m.group(3))])
            rows.append(vals)
        src = f"src/tests/{P}:{l0}-{l0 + text.count(chr(10))}"
        # The tests compare result * 2 sqrt(2) with the table.
        scale = 1 / (2 * math.sqrt(2))
        if gate == "ControlledPhaseShift":
            calls = [([0, 1], 0), ([1, 2], 1)]  # applyControlledPhaseShift / dispatcher sections
            found = re.findall(r"ControlledPhaseShift(?:\(|\",\s*)\{(\d+),\s*(\d+)\},\s*false,\s*\{angles\[(\d)\]\}",
                               text)
            calls = [([int(a), int(b)], int(k)) for a, b, k in found]
            for wires, k in calls:
                out.append({"kind": "gate", "gate": gate, "wires": wires, "params": [angles[k]],
                            "inverse": False, "num_qubits": 3, "init": cvec(H3),
                            "expected": cvec([v * scale for v in rows[k]]), "source": src})
        else:
            for k, a in enumerate(angles):
                out.append({"kind": "gate", "gate": gate, "wires": [k], "params": [a], "inverse": False,
                            "num_qubits": 3, "init": cvec(H3), "expected": cvec([v * scale for v in rows[k]]),
                            "source": src})
    # Rot on |000>: entry 0 = rot[0], entry 1 << (n - i - 1) = rot[2].
    text, l0 = case_text(P, "applyRot\"")
    ang_s, _ = brace_block(text, "angles{")
    ang = numbers(ang_s)
    src = f"src/tests/{P}:{l0}-{l0 + text.count(chr(10))}"
    for i in range(3):
        a = ang[3 * i:3 * i + 3]
        m = rot_mat(*a)
        e = [0j] * 8
        e[0] = m[0]
        e[1 << (3 - i - 1)] = m[2]
        out.append({"kind": "gate", "gate": "Rot", "wires": [i], "params": a, "inverse": False, "num_qubits": 3,
                    "init": cvec([1] + [0] * 7), "expected": cvec(e), "source": src})
    # CRot(0, 1) on |100> (PauliX on wire 0 first): entries 4 and 6.
    text, l0 = case_text(P, "applyCRot\"")
    ang_s, _ = brace_block(text, "angles{")
    a = numbers(ang_s)
    m = rot_mat(*a)
    e = [0j] * 8
    e[1 << 2] = m[0]
    e[(1 << 3) - 2] = m[2]
    out.append({"kind": "gate", "gate": "CRot", "wires": [0, 1], "params": a, "inverse": False, "num_qubits": 3,
                "init": cvec([0, 0, 0, 0, 1, 0, 0, 0]), "expected": cvec(e),
                "source": f"src/tests/{P}:{l0}-{l0 + text.count(chr(10))}"})
    return out


def nonparam_goldens():
    """Basis tables of Test_StateVectorKokkos_NonParam.cpp."""
    N = "Test_StateVectorKokkos_NonParam.cpp"
    out = []
    n = 3
    # Hadamard on |000>: the test checks entry 0 only (1/sqrt 2).
    text, l0 = case_text(N, "applyHadamard\"")
    src = f"src/tests/{N}:{l0}-{l0 + text.count(chr(10))}"
    for w in range(n):
        e = [0j] * 8
        e[0] = 1 / math.sqrt(2)
        out.append({"kind": "gate", "gate": "Hadamard", "wires": [w], "params": [], "inverse": False,
                    "num_qubits": n, "init": cvec([1] + [0] * 7), "expected": cvec(e), "check": [0],
                    "source": src})
    # PauliX on |000>: entry 0 is ZERO, entry 1 << (n - w - 1) is ONE.
    text, l0 = case_text(N, "applyPauliX\"")
    src = f"src/tests/{N}:{l0}-{l0 + text.count(chr(10))}"
    for w in range(n):
        e = [0j] * 8
        e[1 << (n - w - 1)] = UTIL["ONE"]
        out.append({"kind": "gate", "gate": "PauliX", "wires": [w], "params": [], "inverse": False,
                    "num_qubits": n, "init": cvec([1] + [0] * 7), "expected": cvec(e),
                    "check": [0, 1 << (n - w - 1)], "source": src})
    # PauliY, PauliZ, S, T on H(x)H(x)H|000>: expected_results[index] for wire index.
    for gate in ("PauliY", "PauliZ", "S", "T"):
        text, l0 = case_text(N, f"apply{gate}\"")
        env = bind_autos(text, dict(UTIL))
        body, _ = brace_block(text, "expected_results = {")
        rows = [[eval_expr(x, env) for x in split_top(r.strip()[1:-1])] for r in split_top(body)]
        src = f"src/tests/{N}:{l0}-{l0 + text.count(chr(10))}"
        for w in range(n):
            out.append({"kind": "gate", "gate": gate, "wires": [w], "params": [], "inverse": False,
                        "num_qubits": n, "init": cvec(H3), "expected": cvec(rows[w]), "source": src})
    # CNOT(0,1) then CNOT(1,2) on Hadamard(0)|000>: entries 0 and 7 are INVSQRT2.
    text, l0 = case_text(N, "applyCNOT\"")
    calls = re.findall(r"applyCNOT\(\{index - 1, index\}", text)
    assert calls, "CNOT chain not found"
    e = [0j] * 8
    e[0] = e[7] = UTIL["INVSQRT2"]
    init = [0j] * 8
    init[0] = init[4] = UTIL["INVSQRT2"]
    out.append({"kind": "gate_seq", "ops": [["CNOT", [0, 1], [], False], ["CNOT", [1, 2], [], False]],
                "num_qubits": n, "init": cvec(init), "expected": cvec(e), "check": [0, 7],
                "source": f"src/tests/{N}:{l0}-{l0 + text.count(chr(10))}"})
    # SWAP, CZ, Toffoli, CSWAP on |+10> = Hadamard(0) PauliX(1)|000> (the
    # tests' "Check Initial value" section: entries 2 and 6 are INVSQRT2).
    ini = [0j] * 8
    ini[2] = ini[6] = UTIL["INVSQRT2"]
    call = re.compile(r"\.apply(?:(SWAP|CZ|Toffoli|CSWAP)\(|Operation\(\"(\w+)\",\s*)\{([\d,\s]+)\},\s*false\)")
    for gate in ("SWAP", "CZ", "Toffoli", "CSWAP"):
        text, l0 = case_text(N, f"apply{gate}\"")
        env = bind_autos(text, dict(UTIL))
        pos = 0
        while True:
            k = text.find("expected_results = {", pos)
            if k < 0:
                break
            sec_start = text.rfind("SECTION(", 0, k)
            sec_end = text.find("SECTION(", k)
            sec = text[sec_start:sec_end if sec_end > 0 else None]
            senv = bind_autos(sec, dict(env))
            body, _ = brace_block(text, "expected_results = {", k)
            exp = [eval_expr(x, senv) for x in split_top(body)]
            line = text.count("\n", 0, k) + l0
            for m in call.finditer(sec):
                name = m.group(1) or m.group(2)
                wires = [int(x) for x in m.group(3).split(",")]
                out.append({"kind": "gate", "gate": name, "wires": wires, "params": [], "inverse": False,
                            "num_qubits": n, "init": cvec(ini), "expected": cvec(exp),
                            "source": f"src/tests/{N}:{line}"})
            pos = k + 1
    return out


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

    goldens.extend(param_closed_form_goldens())
    goldens.extend(nonparam_goldens())

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
