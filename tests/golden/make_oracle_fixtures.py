"""Writes oracle results at full workload sizes as small JSON fixtures.

The CPU oracle needs minutes for the 24- and 26-qubit workloads, so the GPU
parity tests compare against these stored results instead of rerunning it:
- c3: the 24-qubit strongly entangling circuit (720 parameters, 9
  observables) -- the 9 expectation values and the Jacobian columns of a
  trainable subset (the last 24 parameters plus five from two layers back,
  so the reverse sweep crosses more than two layers);
- c2: the 26-qubit 2000-gate stream -- <H> of the 64-word Hamiltonian, the
  marginal probabilities of wires (0, 13, 25) and 64 amplitudes at seeded
  indices.

Run from the repository root after building the oracle (make -C oracle):
    python tests/golden/make_oracle_fixtures.py [c3] [c2]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from helpers import c2_workload, c3_workload, to_oracle  # noqa: E402

OUT = os.path.join(HERE, "oracle_fixtures.json")

C3_EXTRA = (150, 149, 148, 75, 74)  # distances from the end of the parameter list
C2_PROB_WIRES = [0, 13, 25]
C2_AMP_SEED = 77


def c3_trainable(npar):
    return sorted({npar - d for d in C3_EXTRA} | set(range(npar - 24, npar)))


def c2_amp_indices(n):
    return sorted(int(x) for x in np.random.default_rng(C2_AMP_SEED).choice(2**n, size=64, replace=False))


def make_c3():
    n, ops, obs = c3_workload()
    names, params, wires, inv, mats = ops
    npar = sum(1 for p in params if p)
    tp = c3_trainable(npar)
    oobs = [to_oracle(o) for o in obs]
    t0 = time.time()
    jac = O.adjoint_jacobian(O.zero_state(n), oobs, O.Ops(names, params, wires, inv, mats), tp, True)
    psi = O.apply_ops(O.zero_state(n), O.Ops(names, params, wires, inv, mats))
    ev = [O.expval(psi, o) for o in oobs]
    print(f"c3: {time.time() - t0:.0f} s", flush=True)
    return {"n": n, "trainable": tp, "jacobian": jac.tolist(), "expvals": ev,
            "workload": "helpers.c3_workload() defaults"}


def make_c2():
    n, ops, ham = c2_workload()
    t0 = time.time()
    psi = O.apply_ops(O.zero_state(n), O.Ops(*ops))
    ev = O.expval(psi, to_oracle(ham))
    pr = O.probs(psi, C2_PROB_WIRES)
    idx = c2_amp_indices(n)
    amps = psi[idx]
    print(f"c2: {time.time() - t0:.0f} s", flush=True)
    return {"n": n, "expval_hamiltonian": ev, "prob_wires": C2_PROB_WIRES, "probs": list(map(float, pr)),
            "amp_indices": idx, "amps": [[float(a.real), float(a.imag)] for a in amps],
            "workload": "helpers.c2_workload() defaults"}


def main():
    which = sys.argv[1:] or ["c3", "c2"]
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for w in which:
        data[w] = {"c3": make_c3, "c2": make_c2}[w]()
        data[w]["oracle_threads"] = O.num_threads()
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
