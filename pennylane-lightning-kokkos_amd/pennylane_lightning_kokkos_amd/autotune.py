"""Per-device search of the tile and fusion widths of the adjoint engine.

`tune(n, dtype, n_obs)` times the engine on a synthetic layered circuit of
the caller's size (RZ, RY, RZ on every wire, then a CNOT ring; n_obs PauliZ
observables) for every candidate adjoint tile width, observable-chunk width
and forward tile width, and returns the fastest setting as the environment
variables the engine reads on every call (PLK_TILE_BITS_ADJ, PLK_TILE_NJ,
PLK_TILE_BITS_FWD). With `apply=True` the winner is exported into this
process's environment, so later calls use it; `cache` (a JSON file) keeps
winners per (device, qubits, precision, observables). The settings change
only how commuting operations are grouped and the order of FP64 sums.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Optional, Sequence

import numpy as np

from . import _capi

KNOBS = ("PLK_TILE_BITS_ADJ", "PLK_TILE_NJ", "PLK_TILE_BITS_FWD")


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
         adj_bits: Sequence[int] = (9, 10, 11), chunk: Sequence[int] = (2, 3, 4),
         fwd_bits: Sequence[int] = (9, 10, 11, 12), apply: bool = True,
         cache: Optional[str] = None, device_id: int = 0) -> Dict[str, str]:
    """Fastest (adjoint tile bits, chunk width, forward tile bits) for n
    qubits; candidates outside what the engine accepts at this size are
    skipped (tile bits at most n - 8, LDS budget)."""
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
        timings = {}
        for t in adj_bits:
            if t > n - 8:
                continue
            for nj in chunk:
                os.environ["PLK_TILE_BITS_ADJ"] = str(t)
                os.environ["PLK_TILE_NJ"] = str(nj)
                try:
                    timings[(t, nj)] = _time(
                        lambda: _capi.adjoint_jacobian(state, obs, ops, tp, True), repeats)
                except _capi.PLException:
                    continue
        if timings:
            t, nj = min(timings, key=timings.get)
            best["PLK_TILE_BITS_ADJ"], best["PLK_TILE_NJ"] = str(t), str(nj)
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
