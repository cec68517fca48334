"""Benchmark: adjoint-Jacobian parameters per second at 24 qubits on one MI355X.

Workload (one step): the 24-qubit strongly entangling circuit of 10 layers
(Rot on every wire serialized as RZ, RY, RZ; then CNOT(i, (i+r) mod 24) with
r = (layer mod 23) + 1) = 960 operations and 720 parameters, observables
PauliZ on wires 0-7 plus one 32-term random Pauli Hamiltonian (9 outputs).
A step is the full adjoint Jacobian as the device computes it: forward
application of the circuit to |0...0>, the observables applied to the state,
and the fused backward sweep. Inputs are resident in HBM before timing.

Prints one JSON line. `roofline` refers to the dominant kernel (the
register-tile adjoint pass: one launch is one pass of the backward sweep for
all observables, (1 + 1 + 2 n_obs) S algorithmic bytes): algorithmic bytes
per launch / average launch duration, both measured with HIP events on the
library's stream during the timed steps; `traffic` is the rocprofv3-measured
HBM bytes per launch while the kernel sources match the profiled ones. `cpu_baseline` times the CPU oracle (the reference algorithm restated
in C++/OpenMP) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(REPO, "pennylane-lightning-kokkos_amd"), os.path.join(REPO, "oracle"),
           os.path.join(REPO, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_JSON = os.path.join(REPO, "profiles", "adjoint_traffic.json")


def measured_traffic(family):
    """HBM bytes per launch of the dominant adjoint kernel family measured by rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known size), valid
    only while the kernel sources are the ones that were profiled."""
    import hashlib

    if not os.path.exists(TRAFFIC_JSON):
        return None
    rec = json.load(open(TRAFFIC_JSON))
    if rec.get("family") != family:
        return None
    # The sources the register-tile kernel is compiled from: the bytes a
    # launch moves do not depend on host planning or on other kernels.
    csrc = os.path.join(REPO, "pennylane-lightning-kokkos_amd", "csrc")
    src = [os.path.join(csrc, f) for f in ("kernels_rtile.hip", "device_common.hpp", "plk_internal.hpp")]
    h = hashlib.sha256(b"".join(open(p, "rb").read() for p in src)).hexdigest()[:16]
    return rec["traffic_bytes_per_launch"] if rec.get("source_hash") == h else None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_inputs(capi, n, ops, obs, dtype):
    from helpers import to_capi

    names, params, wires, inv, mats = ops
    s = capi.State(n, dtype)
    cops = capi.Ops(names, params, wires, inv, mats)
    cobs = [to_capi(capi, o) for o in obs]
    npar = sum(1 for p in params if p)
    return s, cops, cobs, list(range(npar))


def run_gpu(capi, n, ops, obs, dtype, steps, warmup, profile=True):
    s, cops, cobs, tp = build_inputs(capi, n, ops, obs, dtype)
    jac = None
    for _ in range(warmup):
        jac = capi.adjoint_jacobian(s, cobs, cops, tp, True)
    s.synchronize()
    if profile:
        capi.profiling_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        jac = capi.adjoint_jacobian(s, cobs, cops, tp, True)
    s.synchronize()
    dt = time.perf_counter() - t0
    stats = {}
    if profile:
        for k in capi.profiling_names():
            stats[k] = capi.profiling_get(k)
        capi.profiling_enable(False)
    return jac, dt, stats, len(tp)


def host_info():
    """CPU model and a measured host memory bandwidth (STREAM-style triad
    a = b + s c over three 1 GiB float64 arrays with torch's CPU threads;
    24 bytes per element counted, write-allocate not counted)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out = {"cpu_model": model, "logical_cpus": os.cpu_count()}
    try:
        import torch

        n = 1 << 27
        a, b, c = torch.empty(n, dtype=torch.float64), torch.rand(n, dtype=torch.float64), torch.rand(n, dtype=torch.float64)
        torch.add(b, c, alpha=0.5, out=a)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            torch.add(b, c, alpha=0.5, out=a)
        dt = time.perf_counter() - t0
        out["host_triad_GBps"] = round(24.0 * n * reps / dt / 1e9, 1)
        out["host_triad_threads"] = torch.get_num_threads()
        del a, b, c
    except Exception as e:  # the baseline is reported without it
        out["host_triad_error"] = str(e)
    return out


def cpu_baseline(n, ops, obs, sample_params):
    """Oracle (reference algorithm, C++/OpenMP) on the same workload, timed in
    three parts: the forward pass (apply every op to |0>), the observable
    application (H_j lambda for every observable) and one adjoint_jacobian
    call whose trainable set is the last `sample_params` parameters (forward +
    observables + reverse sweep until they are consumed). The sweep's cost per
    parameter is that call minus the first two parts, divided by
    `sample_params`; `value` is the full 720-parameter Jacobian rate this
    implies: n_params / (forward + observables + n_params * per-parameter)."""
    import oracle as O
    from helpers import to_oracle

    names, params, wires, inv, mats = ops
    npar = sum(1 for p in params if p)
    tp = list(range(npar - sample_params, npar))
    oops = O.Ops(names, params, wires, inv, mats)
    oobs = [to_oracle(o) for o in obs]
    t0 = time.perf_counter()
    psi = O.apply_ops(O.zero_state(n), oops)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    for o in oobs:
        O.obs_apply(psi, o)
    t_obs = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.adjoint_jacobian(O.zero_state(n), oobs, oops, tp, True)
    t_all = time.perf_counter() - t0
    per_param = max(t_all - t_fwd - t_obs, 1e-9) / sample_params
    full = t_fwd + t_obs + npar * per_param
    info = host_info()
    return {"value": round(npar / full, 3), "unit": "params/s", "cores": O.num_threads(), "kind": "port",
            "sample": (f"oracle (reference algorithm restated in C++/OpenMP) on the same {n}-qubit circuit "
                       f"and {len(obs)} observables: forward pass {t_fwd:.2f} s, observable application "
                       f"{t_obs:.2f} s, adjoint_jacobian with the last {sample_params} of {npar} parameters "
                       f"trainable {t_all:.2f} s; value extrapolates the sweep to all {npar} parameters"),
            "forward_s": round(t_fwd, 3), "observables_s": round(t_obs, 3),
            "sweep_s_per_param": round(per_param, 4),
            "sample_params_per_s": round(sample_params / t_all, 3), **info}


# Bytes each gate of the 26-qubit stream moves when applied alone, as the
# reference's per-gate kernels touch the state: dense or all-diagonal gates
# read and write every amplitude (2 S), controlled gates and SWAP half of
# them (S), CZ a quarter (S / 2).
C2_GATE_BYTES_IN_S = {"Hadamard": 2.0, "RX": 2.0, "RY": 2.0, "RZ": 2.0, "IsingXX": 2.0, "IsingZZ": 2.0,
                      "CNOT": 1.0, "CRX": 1.0, "SWAP": 1.0, "CZ": 0.5}


def _family(stats, names):
    tot = {"launches": 0.0, "total_ms": 0.0, "bytes": 0.0}
    for k in names:
        if k in stats:
            for f in tot:
                tot[f] += stats[k][f]
    return tot


def _collect(capi):
    out = {k: capi.profiling_get(k) for k in capi.profiling_names()}
    capi.profiling_enable(False)
    return out


def run_c2(capi, reps=5):
    """The gate-throughput workload: a 2000-gate stream on a 26-qubit
    complex128 state (1 GiB), <H> of 64 random Pauli words at 26 qubits, and
    the same construction's Hamiltonian at 20 qubits as a CSR matrix
    (device-resident observable; complex128 with int64 indices stored as
    int32, complex64 with the int32 serialisation). Kernel figures come from
    HIP events on the library's stream; GB/s are algorithmic bytes (per kernel
    family: what one launch must read and write) over kernel time."""
    from helpers import c2_workload, pauli_csr, to_capi

    n, ops, ham = c2_workload()
    S = (1 << n) * 16
    s = capi.State(n)
    cops = capi.Ops(*ops)
    s.apply_ops(cops)
    s.synchronize()
    s.reset()
    capi.profiling_enable(True)
    t0 = time.perf_counter()
    s.apply_ops(cops)
    s.synchronize()
    dt = time.perf_counter() - t0
    st = _collect(capi)
    fam = _family(st, ["apply_rtile", "apply_tile", "apply_segment", "apply_dense", "apply_parity"])
    unfused = sum(C2_GATE_BYTES_IN_S[g] for g in ops[0]) * S
    stream = {"gates": len(ops[0]), "ms": round(1e3 * dt, 2), "passes": fam["launches"],
              "gates_per_pass": round(len(ops[0]) / max(1.0, fam["launches"]), 1),
              "kernel_ms": round(fam["total_ms"], 2),
              "kernel_GBps": round(fam["bytes"] / max(fam["total_ms"], 1e-9) / 1e6, 1),
              "kernel_frac_of_peak": round(fam["bytes"] / max(fam["total_ms"], 1e-9) / 1e6 / HBM_PEAK_GBS, 3),
              "unfused_gate_bytes": unfused,
              "unfused_equivalent_GBps": round(unfused / dt / 1e9, 1),
              "gates_per_s": round(len(ops[0]) / dt, 1)}
    log(f"bench: C2 stream {1e3 * dt:.1f} ms, {fam['launches']:.0f} passes, "
        f"{stream['kernel_GBps']} GB/s in pass kernels")
    h = to_capi(capi, ham)
    s.expval(h)
    capi.profiling_enable(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        e = s.expval(h)
    dt = (time.perf_counter() - t0) / reps
    st = _collect(capi)["pauli_expval"]
    k_ms = st["total_ms"] / st["launches"]
    groups = st["bytes"] / st["launches"] / S
    ham_line = {"words": len(ham[1]), "x_mask_groups": int(round(groups)), "call_ms": round(1e3 * dt, 3),
                "kernel_ms": round(k_ms, 3),
                "kernel_GBps": round(st["bytes"] / st["launches"] / (k_ms * 1e-3) / 1e9, 1),
                "per_word_equivalent_GBps": round(len(ham[1]) * S / (k_ms * 1e-3) / 1e9, 1),
                "value": e}
    log(f"bench: C2 <H> 64 words: kernel {k_ms:.3f} ms ({ham_line['x_mask_groups']} X-mask groups)")
    del s
    n20 = 20
    _, _, ham20 = c2_workload(n=n20)
    d, i, p = pauli_csr(n20, ham20, index_dtype=np.int32)
    rows = {}
    for name, dt_ in (("complex128", np.complex128), ("complex64", np.complex64)):
        s20 = capi.State(n20, dt_)
        s20.upload((np.ones(1 << n20) / np.sqrt(1 << n20)).astype(dt_))
        ob = capi.Obs.sparse(d, i, p, list(range(n20)))
        s20.expval(ob)
        capi.profiling_enable(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            s20.expval(ob)
        dtc = (time.perf_counter() - t0) / reps
        st = _collect(capi)["csr_expval"]
        k_ms = st["total_ms"] / st["launches"]
        gbps = st["bytes"] / st["launches"] / (k_ms * 1e-3) / 1e9
        t0 = time.perf_counter()
        s20.expval_csr(d, i, p)
        once = time.perf_counter() - t0
        rows[name] = {"kernel_ms": round(k_ms, 4), "call_ms": round(1e3 * dtc, 3),
                      "kernel_GBps": round(gbps, 1), "kernel_frac_of_peak": round(gbps / HBM_PEAK_GBS, 3),
                      "bytes_per_launch": st["bytes"] / st["launches"],
                      "per_call_upload_ms": round(1e3 * once, 1)}
        log(f"bench: C2 20q CSR {name}: kernel {k_ms:.3f} ms, {gbps:.0f} GB/s")
    csr = {"qubits": n20, "nnz": int(d.size), "index_bytes": 4,
           "bytes_per_nnz": {"complex128": 20, "complex64": 12}, **rows,
           "note": ("kernel bytes = nnz * (value + index) + (rows + 1) * index + state; "
                    "per_call_upload_ms is plk_expval_csr_i32 with host arrays (PCIe upload included)")}
    return {"stream_26q": stream, "hamiltonian_64_words_26q": ham_line, "csr_20q": csr}


def run_c1(capi, reps=50):
    """The plain-method workload (12 qubits, 4 layers of RX, RY, RZ and a
    CNOT ring, 144 parameters, PauliZ on every wire): latency of one full
    adjoint Jacobian (forward pass included) through plk_adjoint_jacobian,
    complex128, median over `reps` calls with new angles each call (handles
    built before the clock, as a training loop would hold them)."""
    from helpers import c1_workload, to_capi

    n, ops, obs = c1_workload()
    rng = np.random.default_rng(20261208)
    s = capi.State(n)
    cobs = [to_capi(capi, o) for o in obs]
    npar = sum(1 for p in ops[1] if p)
    tp = list(range(npar))
    variants = [capi.Ops(ops[0], [[float(rng.uniform(0, 2 * np.pi))] if p else [] for p in ops[1]],
                         ops[2], ops[3], ops[4]) for _ in range(reps)]
    capi.adjoint_jacobian(s, cobs, variants[0], tp, True)
    times = []
    capi.profiling_enable(True)
    for v in variants:
        s.reset()
        t0 = time.perf_counter()
        capi.adjoint_jacobian(s, cobs, v, tp, True)
        times.append(time.perf_counter() - t0)
    st = _collect(capi)
    host = {k: round(v["total_ms"] / v["launches"], 3) for k, v in st.items() if k.startswith("host_adjoint")}
    med = float(np.median(times))
    log(f"bench: C1 12q Jacobian {1e3 * med:.3f} ms (median of {reps})")
    return {"workload": f"{n} qubits, 192 ops, {npar} params, {len(obs)} PauliZ observables, complex128",
            "jacobian_ms_median": round(1e3 * med, 3), "jacobian_ms_min": round(1e3 * min(times), 3),
            "params_per_s": round(npar / med, 1), "reps": reps,
            "host_ms_per_call": host,
            "note": ("structures seen before run as a replayed HIP graph (capture on the second call); "
                     "host_adjoint_plan is the host planning of each call, host_adjoint_graph the table "
                     "upload, replay and result copy")}


def run_excitation(capi, reps=10):
    """Chemistry-style ansatz (20 qubits, Hartree-Fock state, 120
    DoubleExcitation and 60 SingleExcitation gates, a 32-word Hamiltonian and
    4 PauliZ): adjoint Jacobian latency with the single-op step kernels
    (DoubleExcitation touches 2 of its 16 local states, in one launch per
    observable chunk with the Jacobian sums fused) and with them switched off
    (PLK_STEPS=0: the per-state path of a state copy, U on every state,
    G and one inner-product launch per trainable op). Median over `reps`
    calls with new angles, complex128."""
    from helpers import excitation_workload, to_capi

    n, ops, obs = excitation_workload()
    cobs = [to_capi(capi, o) for o in obs]
    npar = sum(1 for p in ops[1] if p)
    tp = list(range(npar))
    rng = np.random.default_rng(20261210)
    variants = [capi.Ops(ops[0], [[float(rng.uniform(0, 2 * np.pi))] if p else [] for p in ops[1]],
                         ops[2], ops[3], ops[4]) for _ in range(reps)]
    s = capi.State(n)
    out = {"workload": f"{n} qubits, Hartree-Fock state, {ops[0].count('DoubleExcitation')} "
                       f"DoubleExcitation + {ops[0].count('SingleExcitation')} SingleExcitation, "
                       f"{npar} params, 32-word Hamiltonian + 4 PauliZ, complex128"}
    jacs = {}
    old = os.environ.get("PLK_STEPS")
    try:
        for mode, key in (("1", "step_kernels"), ("0", "per_state_path")):
            os.environ["PLK_STEPS"] = mode
            s.reset()
            capi.adjoint_jacobian(s, cobs, variants[0], tp, True)
            times = []
            for v in variants:
                s.reset()
                t0 = time.perf_counter()
                jacs[key] = capi.adjoint_jacobian(s, cobs, v, tp, True)
                times.append(time.perf_counter() - t0)
            med = float(np.median(times))
            out[key] = {"jacobian_ms_median": round(1e3 * med, 3), "params_per_s": round(npar / med, 1)}
    finally:
        if old is None:
            os.environ.pop("PLK_STEPS", None)
        else:
            os.environ["PLK_STEPS"] = old
    out["speedup"] = round(out["per_state_path"]["jacobian_ms_median"] /
                           out["step_kernels"]["jacobian_ms_median"], 2)
    out["max_abs_diff"] = float(np.max(np.abs(jacs["step_kernels"] - jacs["per_state_path"])))
    log(f"bench: excitation ansatz {out['step_kernels']['jacobian_ms_median']} ms "
        f"(per-state path {out['per_state_path']['jacobian_ms_median']} ms)")
    return out


MFMA_FP32_PEAK_TFLOPS = 157.3


def run_c5(capi, n=32, layers=40, reference=True):
    """Large-state fused form: the 32-qubit complex64 circuit of 40 layers of
    Haar SU(2) on every wire and CZ bricks, fused into 5-qubit 32x32 blocks
    applied on the matrix cores; the same circuit on complex128 through the
    plain path is the reference (error reported on the TFIM energy, the norm
    and 4096 sampled amplitudes), and complex64 through the plain path is
    timed beside it. Per block pass: 2 S bytes and 8 * 32 * 2^n flops."""
    from helpers import c5_workload, to_capi

    n, ops, ham = c5_workload(n=n, layers=layers)
    cops = capi.Ops(*ops)
    h = to_capi(capi, ham)
    out = {"workload": f"{n} qubits, {layers} layers of Haar SU(2) on every wire + CZ brick, "
                       f"{len(ops[0])} ops, TFIM observable", "qubits": n, "ops": len(ops[0])}
    a = capi.State(n, np.complex64)
    capi.profiling_enable(True)
    t0 = time.perf_counter()
    nb = a.apply_ops_blocks(cops)
    a.synchronize()
    dt = time.perf_counter() - t0
    st = _collect(capi)["apply_block5"]
    k_ms = st["total_ms"] / st["launches"]
    S = (1 << n) * 8
    flops = 8.0 * 32 * (1 << n)
    gbps = 2.0 * S / (k_ms * 1e-3) / 1e9
    tfl = flops / (k_ms * 1e-3) / 1e12
    out["blocks_complex64"] = {"ms": round(1e3 * dt, 1), "blocks": nb, "ops_per_block": round(len(ops[0]) / nb, 2),
                               "kernel_ms_per_block": round(k_ms, 3), "kernel_GBps": round(gbps, 1),
                               "hbm_frac": round(gbps / HBM_PEAK_GBS, 3), "mfma_TFLOPs": round(tfl, 1),
                               "mfma_frac": round(tfl / MFMA_FP32_PEAK_TFLOPS, 3)}
    log(f"bench: C5 blocks {1e3 * dt:.0f} ms, {nb} blocks, {k_ms:.2f} ms/block, {gbps:.0f} GB/s, {tfl:.1f} TFLOP/s")
    if not reference:
        return out
    e64 = a.expval(h)
    idx = np.random.default_rng(32).integers(0, 1 << n, size=4096)
    amps64 = _amplitudes(capi, a, idx)
    norm64 = float(np.sqrt(a.probs([0]).sum()))
    del a
    b = capi.State(n, np.complex128)
    t0 = time.perf_counter()
    b.apply_ops(cops)
    b.synchronize()
    dt128 = time.perf_counter() - t0
    e128 = b.expval(h)
    amps128 = _amplitudes(capi, b, idx)
    del b
    c = capi.State(n, np.complex64)
    t0 = time.perf_counter()
    c.apply_ops(cops)
    c.synchronize()
    dtp = time.perf_counter() - t0
    del c
    out["plain_complex128_ms"] = round(1e3 * dt128, 1)
    out["plain_complex64_ms"] = round(1e3 * dtp, 1)
    out["error_vs_complex128"] = {"tfim_energy": abs(e64 - e128), "energy": e128,
                                  "max_abs_amplitude_4096_samples": float(np.max(np.abs(amps64 - amps128))),
                                  "norm_minus_1": norm64 - 1.0}
    log(f"bench: C5 complex128 plain {1e3 * dt128:.0f} ms, complex64 plain {1e3 * dtp:.0f} ms, "
        f"|dE| = {abs(e64 - e128):.2e}")
    return out


def _amplitudes(capi, s, idx):
    """Amplitudes at the given indices, copied from the device one by one
    (a 32-qubit state is 32-64 GiB: it is never downloaded whole)."""
    import ctypes

    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    s.synchronize()
    base = s.device_ptr()
    item = 8 if s.dtype == np.complex64 else 16
    out = np.zeros(len(idx), dtype=s.dtype)
    for k, i in enumerate(idx):
        rc = hip.hipMemcpy(out[k:].ctypes.data, base + int(i) * item, item, 2)  # device to host
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
    return out.astype(np.complex128)


def copy_ceiling(capi, n=24, reps=10):
    """Device-to-device copies of a complex128 state (hipMemcpyAsync): read +
    written bytes over time, the practical HBM ceiling of a streaming pass."""
    a, b = capi.State(n), capi.State(n)
    b.copy_from(a)
    capi.profiling_enable(True)
    for _ in range(reps):
        b.copy_from(a)
    b.synchronize()
    st = _collect(capi)["copy"]
    return round(st["bytes"] / (st["total_ms"] * 1e-3) / 1e9, 1)


def run_batched(capi, count, seq_count):
    """The 4096-circuit batched workload: n uniform in [6, 16], 2-8 layers of
    RY and RZ on every wire plus a CNOT chain, observable sum_i Z_i, every
    parameter trainable. Timed: one plk_adjoint_batched call in complex64
    (FP64 accumulation) and in complex128, each after one untimed call, and
    the plain per-circuit complex128 path (plk_adjoint_jacobian with
    apply_operations, states preallocated per qubit count) over the first
    seq_count circuits. Handles are built before the clocks start."""
    from helpers import c4_workload

    circuits = c4_workload(count=count)
    nq = [c[0] for c in circuits]
    ops = [capi.Ops(*c[1]) for c in circuits]
    # A second angle set for the same circuits: a training step's next call
    # (same structure, new parameters), built before the clocks start.
    rng = np.random.default_rng(20261206)
    ops_next = [capi.Ops(c[1][0], [[float(rng.uniform(0, 2 * np.pi))] if p else [] for p in c[1][1]],
                         c[1][2], c[1][3], c[1][4]) for c in circuits]
    obs = [capi.Obs.hamiltonian(np.ones(c[0]), [capi.Obs.named("PauliZ", [w]) for w in range(c[0])])
           for c in circuits]
    params = sum(sum(1 for p in c[1][1] if p) for c in circuits)
    out = {"workload": (f"{count} circuits, n uniform in [6,16], 2-8 layers of RY and RZ on every "
                        f"wire then CNOT chain, observable sum_i Z_i, all parameters trainable"),
           "circuits": count, "params": params,
           "timing": ("first_call_ms plans the batch (host) and uploads its tables; call_ms is the "
                      "next call with new angles for every parameter (resident plan, angles "
                      "uploaded and patched on the device); both include the Python wrapper")}
    rows = {}
    for name, dt in (("complex64", np.complex64), ("complex128", np.complex128)):
        capi.profiling_enable(True)
        t0 = time.perf_counter()
        capi.adjoint_batched(nq, ops, obs, None, dtype=dt)
        dt_first = time.perf_counter() - t0
        first = _collect(capi)
        capi.profiling_enable(True)
        t0 = time.perf_counter()
        ev, r = capi.adjoint_batched(nq, ops_next, obs, None, dtype=dt)
        dt_call = time.perf_counter() - t0
        st = _collect(capi)
        k = st["adjoint_batch"]
        rows[name] = (ev, r)
        host = {key: round(v["total_ms"], 2) for key, v in st.items() if key.startswith("host_batch")}
        host_first = {key: round(v["total_ms"], 2) for key, v in first.items()
                      if key.startswith("host_batch") and not key.endswith("bytes")}
        out[name] = {"call_ms": round(1e3 * dt_call, 2), "kernel_ms": round(k["total_ms"], 3),
                     "patch_kernel_ms": round(st.get("batch_patch", {"total_ms": 0.0})["total_ms"], 3),
                     "call_over_kernel": round(1e3 * dt_call / k["total_ms"], 2),
                     "first_call_ms": round(1e3 * dt_first, 1),
                     "params_per_s_call": round(params / dt_call, 1),
                     "params_per_s_kernel": round(params / (k["total_ms"] * 1e-3), 1),
                     "host_ms": host, "host_ms_first_call": host_first,
                     "table_bytes": first.get("host_batch_table_bytes", {"bytes": 0.0})["bytes"]}
        log(f"bench: batched {name}: first call {1e3 * dt_first:.1f} ms, call {1e3 * dt_call:.1f} ms, "
            f"kernel {k['total_ms']:.2f} ms")
    states = {}
    seq = circuits[:seq_count]
    seq_params = sum(sum(1 for p in c[1][1] if p) for c in seq)
    for n in sorted(set(c[0] for c in seq)):
        states[n] = capi.State(n, np.complex128)
    jacs = []
    t0 = time.perf_counter()
    for i, c in enumerate(seq):
        s = states[c[0]]
        s.reset()
        npar = sum(1 for p in c[1][1] if p)
        jacs.append(capi.adjoint_jacobian(s, [obs[i]], ops_next[i], list(range(npar)), True)[0])
    dt_seq = time.perf_counter() - t0
    out["sequential_complex128"] = {"circuits": len(seq), "ms": round(1e3 * dt_seq, 1),
                                    "params_per_s": round(seq_params / dt_seq, 1)}
    ev64, r64 = rows["complex64"]
    ev128, r128 = rows["complex128"]
    out["max_abs_err_complex64_vs_per_circuit_complex128"] = float(
        max(np.max(np.abs(r64[i] - jacs[i])) for i in range(len(seq))))
    out["max_abs_err_complex128_vs_per_circuit_complex128"] = float(
        max(np.max(np.abs(r128[i] - jacs[i])) for i in range(len(seq))))
    out["speedup_complex64_call_vs_sequential"] = round(
        out["complex64"]["params_per_s_call"] / out["sequential_complex128"]["params_per_s"], 1)
    log(f"bench: sequential per-circuit complex128: {len(seq)} circuits {dt_seq:.2f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=24)
    ap.add_argument("--c64-steps", type=int, default=3)
    ap.add_argument("--c1", type=int, default=1, help="time the 12-qubit Jacobian latency (0 skips it)")
    ap.add_argument("--c2", type=int, default=1, help="run the gate-throughput workload (0 skips it)")
    ap.add_argument("--excitation", type=int, default=1,
                    help="time the 20-qubit DoubleExcitation ansatz (0 skips it)")
    ap.add_argument("--c5", type=int, default=1, help="run the 32-qubit fused-block workload (0 skips it)")
    ap.add_argument("--c5-reference", type=int, default=1,
                    help="also run the complex128 / complex64 plain paths of that workload (0: blocks only)")
    ap.add_argument("--batched", type=int, default=4096,
                    help="circuits of the batched workload (0 skips it)")
    ap.add_argument("--batched-seq", type=int, default=4096,
                    help="circuits timed through the per-circuit path for comparison")
    ap.add_argument("--probe-copy", action="store_true",
                    help="only copy one 24-qubit complex128 state to another 10 times "
                         "(256 MiB read + 256 MiB written per copy): the calibration "
                         "reference for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path "
              "(sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from helpers import c3_workload
    from pennylane_lightning_kokkos_amd import _capi as capi

    capi.load()
    if args.probe_copy:
        a, b = capi.State(24), capi.State(24)
        for _ in range(10):
            b.copy_from(a)
        b.synchronize()
        print(json.dumps({"probe": "copy", "bytes_read_per_copy": 2**28, "bytes_written_per_copy": 2**28}))
        return
    n, ops, obs = c3_workload(n=args.qubits)
    log(f"bench: {n} qubits, {len(ops[0])} ops, {len(obs)} observables; complex128 warmup")
    jac, dt, stats, ntp = run_gpu(capi, n, ops, obs, np.complex128, args.steps, args.warmup)
    ms = 1e3 * dt / args.steps
    value = ntp * args.steps / dt
    log(f"bench: complex128 {ms:.1f} ms/step, {value:.1f} params/s")

    ceiling = copy_ceiling(capi)
    log(f"bench: state copy {ceiling} GB/s")
    fams = [k for k in stats if k.startswith("adjoint_")]
    dom = max(fams, key=lambda k: stats[k]["total_ms"]) if fams else None
    seg = stats.get(dom, {"launches": 0, "total_ms": 0, "bytes": 0})
    roof = None
    if seg["launches"]:
        avg_ms = seg["total_ms"] / seg["launches"]
        avg_bytes = seg["bytes"] / seg["launches"]
        achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                "kernel": {"adjoint_rtile": "rtile_kernel (register-tile pass, adjoint)",
                           "adjoint_tile": "tile_kernel (LDS tile pass, adjoint)",
                           "adjoint_segment": "seg_kernel (register segment, adjoint)"}.get(dom, dom),
                "avg_launch_ms": round(avg_ms, 4), "avg_launch_bytes": avg_bytes,
                "launches_per_step": seg["launches"] / args.steps,
                "copy_ceiling_GBps": ceiling, "frac_of_copy_ceiling": round(achieved / ceiling, 4),
                "note": ("bytes per launch are fixed by the pass (every state read and written "
                         "once), so packing more operations into a pass lowers GB/s while it "
                         "cuts the passes, and the bytes, of a Jacobian")}
    # Per family: algorithmic bytes (what a launch must read and write; the
    # Pauli-sum kernels count every psi[i ^ x] gather as a read, many of which
    # hit in cache) over HIP-event time. Host scopes carry no bytes.
    breakdown = {k: {"ms_per_step": round(v["total_ms"] / args.steps, 3),
                     "launches_per_step": v["launches"] / args.steps,
                     "algorithmic_GBps": round(v["bytes"] / max(v["total_ms"], 1e-9) / 1e6, 1)}
                 for k, v in stats.items()}

    # Sweep-level accounting: the unfused algorithm moves 2 S (1 + n_obs) bytes
    # per trainable rotation and S (1 + n_obs) per CNOT (every state read and
    # written once per gate). The fused passes move far fewer bytes;
    # unfused_equivalent_GBps states the bandwidth a one-gate-per-pass sweep
    # would need to match this sweep's time (it is not a measured bandwidth).
    S = (2 ** n) * 16
    n_rot = sum(1 for p in ops[1] if p)
    n_other = len(ops[0]) - n_rot
    unfused = (2 * n_rot + n_other) * S * (1 + len(obs))
    adj_ms = sum(v["total_ms"] for k, v in stats.items()
                 if k.startswith("adjoint_") or k.startswith("reduce")) / args.steps
    fwd_ms = sum(v["total_ms"] for k, v in stats.items() if k.startswith("apply_")) / args.steps
    # One register-tile launch is one pass for all observables (its chunks
    # run side by side).
    passes = seg["launches"] / args.steps if seg["launches"] else None
    extra = {"sweep": {"unfused_algorithmic_bytes": unfused, "adjoint_ms_per_step": round(adj_ms, 3),
                       "forward_ms_per_step": round(fwd_ms, 3),
                       "unfused_equivalent_GBps": round(unfused / (adj_ms * 1e-3) / 1e9, 1),
                       "ops_per_adjoint_pass": round(len(ops[0]) / passes, 1) if passes else None}}
    if args.c64_steps > 0:
        log("bench: complex64 (FP64 reductions)")
        jac64, dt64, _, _ = run_gpu(capi, n, ops, obs, np.complex64, args.c64_steps, 1, profile=False)
        extra["complex64"] = {"params_per_s": round(ntp * args.c64_steps / dt64, 1),
                              "ms_per_step": round(1e3 * dt64 / args.c64_steps, 2),
                              "max_abs_diff_vs_complex128": float(np.max(np.abs(jac64 - jac)))}

    if args.c1:
        extra["c1"] = run_c1(capi)

    if args.excitation:
        extra["excitation"] = run_excitation(capi)

    if args.c2:
        log("bench: gate-throughput workload (26-qubit stream, 64-word <H>, 20-qubit CSR)")
        extra["c2"] = run_c2(capi)

    if args.c5:
        log("bench: large-state fused form (32 qubits, MFMA blocks)")
        extra["c5"] = run_c5(capi, reference=bool(args.c5_reference))

    if args.batched > 0:
        log(f"bench: batched workload ({args.batched} circuits)")
        extra["batched"] = run_batched(capi, args.batched, min(args.batched_seq, args.batched))

    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU oracle baseline (last {args.cpu_sample} parameters)")
        cpu = cpu_baseline(n, ops, obs, args.cpu_sample)
        log(f"bench: CPU {cpu['value']:.2f} params/s on {cpu['cores']} threads")

    out = {
        "metric": "adjoint Jacobian params/s @24q (complex128, FP64 reductions)",
        "value": round(value, 2),
        "unit": "params/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded circuit parameters and Hamiltonian)",
        "config": {"workload": f"adjoint Jacobian, {n}-qubit strongly entangling circuit, 10 layers, "
                               f"960 ops, 720 params, 8 PauliZ + 32-term Pauli Hamiltonian",
                   "qubits": n, "params": ntp, "observables": len(obs), "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": breakdown,
        **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
