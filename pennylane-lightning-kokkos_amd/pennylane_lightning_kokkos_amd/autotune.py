"""Per-device search of the pass kind, tile width and chunk width of the
adjoint engine.

`tune(n, dtype, n_obs)` times full adjoint Jacobians on a synthetic layered
circuit of the caller's size (RZ, RY, RZ on every wire, then a CNOT ring;
n_obs PauliZ observables) for every candidate: register-tile passes with
each observable-chunk width (PLK_RTILE=1, PLK_RTILE_NJ), and LDS tile passes
with each adjoint tile width and chunk width (PLK_RTILE=0,
PLK_TILE_BITS_ADJ, PLK_TILE_NJ). When the LDS passes win, the forward tile
width (PLK_TILE_BITS_FWD) is searched as well. The fastest setting is
returned as the environment variables the engine reads on every call; with
`apply=True` it is exported into this process's environment, so later calls
use it; `cache` (a JSON file) keeps winners per (device, qubits, precision,
observables). The settings change only how commuting operations are grouped
and the order of FP64 sums.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Optional, Sequence

import numpy as np

from . import _capi

KNOBS = ("PLK_RTILE", "PLK_RTILE_NJ", "PLK_TILE_BITS_ADJ", "PLK_TILE_NJ", "PLK_TILE_BITS_FWD")

# Register bits of a register tile (complex128, complex64); a register-tile
# pass needs at least that many qubits plus the six lane bits.
_RT_REGISTER_BITS = {np.dtype(np.complex128): 3, np.dtype(np.complex64): 4}


def _circuit(n: int, layers: int, seed: int):
    rng = np.random.default_rng(seed)
    names, params, wires = [], [], []
    for _ in range(layers):
        for w in range(n):
            for g in ("RZ", "RY", "RZ"):
                names.append(g)
                params.append([float(rng.uniform(0, 2 * np.pi))])
                wires.append([w])
        for w in range(n):
            names.append("CNOT")
            params.append([])
            wires.append([w, (w + 1) % n])
    return names, params, wires


def _time(fn, repeats: int) -> float:
    fn()
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t0)
    return best


def _key(n, dtype, n_obs) -> str:
    return f"{_capi.print_configuration().splitlines()[0]}|n={n}|{np.dtype(dtype).name}|obs={n_obs}"


def tune(n: int, dtype=np.complex128, n_obs: int = 9, layers: int = 4, repeats: int = 2,
         rt_chunk: Sequence[int] = (2, 3, 4), adj_bits: Sequence[int] = (9, 10, 11),
         chunk: Sequence[int] = (2, 3, 4), fwd_bits: Sequence[int] = (9, 10, 11, 12),
         apply: bool = True, cache: Optional[str] = None, device_id: int = 0) -> Dict[str, str]:
    """Fastest setting for n qubits; candidates outside what the engine
    accepts at this size are skipped (register tiles need R + 6 qubits, LDS
    tile bits at most n - 8, LDS budget)."""
    if cache and os.path.exists(cache):
        hit = json.load(open(cache)).get(_key(n, dtype, n_obs))
        if hit:
            if apply:
                os.environ.update(hit)
            return dict(hit)
    names, params, wires = _circuit(n, layers, seed=n)
    ops = _capi.Ops(names, params, wires, [False] * len(names), [[] for _ in names])
    obs = [_capi.Obs.named("PauliZ", [w % n]) for w in range(n_obs)]
    tp = list(range(sum(1 for p in params if p)))
    state = _capi.State(n, dtype, device_id)
    saved = {k: os.environ.get(k) for k in KNOBS}
    best: Dict[str, str] = {}
    try:
        for k in KNOBS:
            os.environ.pop(k, None)
        cands = []
        if n >= _RT_REGISTER_BITS[np.dtype(dtype)] + 6:
            cands += [{"PLK_RTILE": "1", "PLK_RTILE_NJ": str(nj)} for nj in rt_chunk]
        cands += [{"PLK_RTILE": "0", "PLK_TILE_BITS_ADJ": str(t), "PLK_TILE_NJ": str(nj)}
                  for t in adj_bits if t <= n - 8 for nj in chunk]
        timings = []
        for c in cands:
            os.environ.update(c)
            try:
                timings.append((_time(lambda: _capi.adjoint_jacobian(state, obs, ops, tp, True),
                                      repeats), c))
            except _capi.PLException:
                pass
            for k in c:
                os.environ.pop(k, None)
        if timings:
            best = dict(min(timings, key=lambda tc: tc[0])[1])
        if best.get("PLK_RTILE") == "0":
            os.environ.update(best)
            ftimes = {}
            for t in fwd_bits:
                if t > n - 8:
                    continue
                os.environ["PLK_TILE_BITS_FWD"] = str(t)

                def fwd():
                    state.reset()
                    state.apply_ops(ops)
                    state.synchronize()

                try:
                    ftimes[t] = _time(fwd, repeats)
                except _capi.PLException:
                    continue
            if ftimes:
                best["PLK_TILE_BITS_FWD"] = str(min(ftimes, key=ftimes.get))
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    if apply:
        os.environ.update(best)
    if cache:
        data = json.load(open(cache)) if os.path.exists(cache) else {}
        data[_key(n, dtype, n_obs)] = best
        with open(cache, "w") as f:
            json.dump(data, f, indent=1)
    return best
