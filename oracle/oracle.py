"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle (``oracle.cpp``) is a plain C++/OpenMP restatement of the
PennyLane-Lightning-Kokkos algorithms on the adjoint-Jacobian path. Only the
tests, the smoke check and the benchmark's CPU-baseline leg import this
module, as the checker or as the timed CPU baseline. Nothing in the product
package imports it.

Every state here is a numpy ``complex128`` vector of length ``2**n``; wire
``w`` is bit ``n-1-w`` of the amplitude index
(``pennylane_lightning_kokkos/src/simulator/GateFunctors.hpp:32``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_obs_named.restype = ctypes.c_void_p
        L.orc_obs_hermitian.restype = ctypes.c_void_p
        L.orc_obs_tensor.restype = ctypes.c_void_p
        L.orc_obs_hamiltonian.restype = ctypes.c_void_p
        L.orc_obs_sparse.restype = ctypes.c_void_p
        L.orc_ops_create.restype = ctypes.c_void_p
        L.orc_obs_free.argtypes = [ctypes.c_void_p]
        L.orc_ops_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc: int):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_c_double_p)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_c_int64_p)


def _as_i64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int64).reshape(-1))


def _as_c128(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128).reshape(-1))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ----------------------------------------------------------------------------
# State-level operations
# ----------------------------------------------------------------------------
def zero_state(n: int) -> np.ndarray:
    sv = np.zeros(2**n, dtype=np.complex128)
    sv[0] = 1.0
    return sv


def apply(sv: np.ndarray, name: str, wires: Sequence[int], inverse: bool = False,
          params: Sequence[float] = (), matrix=None) -> np.ndarray:
    """StateVectorKokkos::applyOperation (StateVectorKokkos.hpp:585-600), in place."""
    assert sv.dtype == np.complex128 and sv.flags.c_contiguous
    n = int(np.log2(sv.size))
    w = _as_i64(wires)
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
    m = None if matrix is None else _as_c128(matrix)
    _check(lib().orc_apply(_dp(sv), n, name.encode(), _ip(w), len(w), int(bool(inverse)),
                           _dp(p), len(p), None if m is None else _dp(m),
                           ctypes.c_int64(0 if m is None else m.size)))
    return sv


def apply_matrix(sv: np.ndarray, matrix, wires: Sequence[int], inverse: bool = False):
    n = int(np.log2(sv.size))
    w = _as_i64(wires)
    m = _as_c128(matrix)
    _check(lib().orc_apply_matrix(_dp(sv), n, _dp(m), _ip(w), len(w), int(bool(inverse))))
    return sv


def apply_generator(sv: np.ndarray, name: str, wires: Sequence[int]) -> float:
    """StateVectorKokkos::applyGenerator (StateVectorKokkos.hpp:687-695), in place."""
    n = int(np.log2(sv.size))
    w = _as_i64(wires)
    s = ctypes.c_double()
    _check(lib().orc_apply_generator(_dp(sv), n, name.encode(), _ip(w), len(w), ctypes.byref(s)))
    return s.value


def gate_info(name: str):
    npar = ctypes.c_int()
    nw = ctypes.c_int()
    _check(lib().orc_gate_info(name.encode(), ctypes.byref(npar), ctypes.byref(nw)))
    return npar.value, nw.value


def gate_matrix(name: str, nwires: int, params=(), inverse=False) -> np.ndarray:
    dim = 2**nwires
    out = np.zeros(dim * dim, dtype=np.complex128)
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
    _check(lib().orc_gate_matrix(name.encode(), nwires, int(bool(inverse)), _dp(p), len(p), _dp(out)))
    return out.reshape(dim, dim)


def generator_matrix(name: str, nwires: int):
    dim = 2**nwires
    out = np.zeros(dim * dim, dtype=np.complex128)
    s = ctypes.c_double()
    _check(lib().orc_generator_matrix(name.encode(), nwires, _dp(out), ctypes.byref(s)))
    return out.reshape(dim, dim), s.value


def inner(x: np.ndarray, y: np.ndarray):
    re = ctypes.c_double()
    im = ctypes.c_double()
    _check(lib().orc_inner(_dp(x), _dp(y), ctypes.c_int64(x.size), ctypes.byref(re), ctypes.byref(im)))
    return complex(re.value, im.value)


def spmv(x: np.ndarray, data, indices, indptr) -> np.ndarray:
    d = _as_c128(data)
    i = _as_i64(indices)
    p = _as_i64(indptr)
    y = np.zeros(len(p) - 1, dtype=np.complex128)
    _check(lib().orc_spmv(_dp(x), _dp(y), _dp(d), _ip(i), _ip(p), ctypes.c_int64(len(p) - 1),
                          ctypes.c_int64(d.size)))
    return y


def expval_csr(sv: np.ndarray, data, indices, indptr) -> float:
    d = _as_c128(data)
    i = _as_i64(indices)
    p = _as_i64(indptr)
    out = ctypes.c_double()
    n = int(np.log2(sv.size))
    _check(lib().orc_expval_csr(_dp(sv), n, _dp(d), _ip(i), _ip(p), ctypes.c_int64(len(p) - 1),
                                ctypes.byref(out)))
    return out.value


def probs(sv: np.ndarray, wires=()) -> np.ndarray:
    """MeasuresKokkos::probs(wires): marginals, index bits in the given wire order."""
    n = int(np.log2(sv.size))
    w = _as_i64(list(wires))
    out = np.zeros(2 ** (w.size if w.size else n))
    x = np.ascontiguousarray(sv, dtype=np.complex128)
    _check(lib().orc_probs(_dp(x), n, _ip(w), int(w.size), _dp(out)))
    return out


def sample_indices(sv: np.ndarray, uniforms) -> np.ndarray:
    """Basis indices that generate_samples selects for the given uniforms."""
    n = int(np.log2(sv.size))
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    out = np.zeros(max(1, u.size), dtype=np.int64)
    x = np.ascontiguousarray(sv, dtype=np.complex128)
    _check(lib().orc_sample_indices(_dp(x), n, _dp(u), ctypes.c_int64(u.size), _ip(out)))
    return out[: u.size]


# ----------------------------------------------------------------------------
# Observables (ObservablesKokkos.hpp:19-517) as plain Python descriptions;
# converted to oracle handles on demand.
# ----------------------------------------------------------------------------
@dataclass
class Named:
    name: str
    wires: List[int]


@dataclass
class Hermitian:
    matrix: np.ndarray
    wires: List[int]


@dataclass
class Tensor:
    factors: list


@dataclass
class Hamiltonian:
    coeffs: List[float]
    terms: list


@dataclass
class Sparse:
    data: np.ndarray
    indices: np.ndarray
    indptr: np.ndarray
    wires: List[int]


class _Handle:
    def __init__(self, ptr):
        if not ptr:
            raise OracleError(lib().orc_last_error().decode())
        self.ptr = ctypes.c_void_p(ptr)

    def __del__(self):
        try:
            lib().orc_obs_free(self.ptr)
        except Exception:  # pragma: no cover - interpreter shutdown
            pass


def _obs_handle(ob) -> _Handle:
    L = lib()
    if isinstance(ob, Named):
        w = _as_i64(ob.wires)
        return _Handle(L.orc_obs_named(ob.name.encode(), _ip(w), len(w)))
    if isinstance(ob, Hermitian):
        w = _as_i64(ob.wires)
        m = _as_c128(ob.matrix)
        return _Handle(L.orc_obs_hermitian(_dp(m), _ip(w), len(w)))
    if isinstance(ob, Tensor):
        hs = [_obs_handle(f) for f in ob.factors]
        arr = (ctypes.c_void_p * len(hs))(*[h.ptr.value for h in hs])
        return _Handle(L.orc_obs_tensor(arr, len(hs)))
    if isinstance(ob, Hamiltonian):
        hs = [_obs_handle(t) for t in ob.terms]
        arr = (ctypes.c_void_p * len(hs))(*[h.ptr.value for h in hs])
        c = np.ascontiguousarray(np.asarray(ob.coeffs, dtype=np.float64))
        return _Handle(L.orc_obs_hamiltonian(_dp(c), arr, len(hs)))
    if isinstance(ob, Sparse):
        d = _as_c128(ob.data)
        i = _as_i64(ob.indices)
        p = _as_i64(ob.indptr)
        w = _as_i64(ob.wires)
        return _Handle(L.orc_obs_sparse(_dp(d), _ip(i), _ip(p), ctypes.c_int64(d.size),
                                        ctypes.c_int64(len(p) - 1), _ip(w), len(w)))
    raise TypeError(f"unknown observable {ob!r}")


def obs_apply(sv: np.ndarray, ob) -> np.ndarray:
    h = _obs_handle(ob)
    n = int(np.log2(sv.size))
    _check(lib().orc_obs_apply(_dp(sv), n, h.ptr))
    return sv


def expval(sv: np.ndarray, ob) -> float:
    """MeasuresKokkos::expval(ob) (MeasuresKokkos.hpp:354-360)."""
    h = _obs_handle(ob)
    out = ctypes.c_double()
    n = int(np.log2(sv.size))
    _check(lib().orc_expval_obs(_dp(sv), n, h.ptr, ctypes.byref(out)))
    return out.value


# ----------------------------------------------------------------------------
# Operations + adjoint Jacobian (AdjointDiffKokkos.hpp:17-173, 404-478)
# ----------------------------------------------------------------------------
@dataclass
class Ops:
    names: List[str]
    params: List[List[float]]
    wires: List[List[int]]
    inverses: List[bool]
    matrices: list = field(default=None)

    def __post_init__(self):
        if self.matrices is None:
            self.matrices = [[] for _ in self.names]

    def num_par_ops(self) -> int:
        return sum(1 for p in self.params if len(p) > 0)


class _OpsHandle:
    def __init__(self, ops: Ops):
        L = lib()
        n = len(ops.names)
        names = (ctypes.c_char_p * n)(*[s.encode() for s in ops.names])
        self._keep = names
        poff = np.zeros(n + 1, dtype=np.int64)
        woff = np.zeros(n + 1, dtype=np.int64)
        moff = np.zeros(n + 1, dtype=np.int64)
        for i in range(n):
            poff[i + 1] = poff[i] + len(ops.params[i])
            woff[i + 1] = woff[i] + len(ops.wires[i])
            moff[i + 1] = moff[i] + np.asarray(ops.matrices[i]).size
        params = np.ascontiguousarray(
            np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1) for p in ops.params] + [np.zeros(1)]))
        wires = np.ascontiguousarray(
            np.concatenate([np.asarray(w, dtype=np.int64).reshape(-1) for w in ops.wires] + [np.zeros(1, np.int64)]))
        mats = np.ascontiguousarray(
            np.concatenate([np.asarray(m, dtype=np.complex128).reshape(-1) for m in ops.matrices]
                           + [np.zeros(1, np.complex128)]))
        inv = (ctypes.c_int * n)(*[int(bool(b)) for b in ops.inverses])
        ptr = L.orc_ops_create(n, names, _dp(params), _ip(poff), _ip(wires), _ip(woff), inv,
                               _dp(mats.view(np.float64)), _ip(moff))
        if not ptr:
            raise OracleError(L.orc_last_error().decode())
        self.ptr = ctypes.c_void_p(ptr)

    def __del__(self):
        try:
            lib().orc_ops_free(self.ptr)
        except Exception:  # pragma: no cover
            pass


def apply_ops(sv: np.ndarray, ops: Ops) -> np.ndarray:
    for i, name in enumerate(ops.names):
        m = ops.matrices[i]
        apply(sv, name, ops.wires[i], ops.inverses[i], ops.params[i],
              None if np.asarray(m).size == 0 else m)
    return sv


def adjoint_jacobian(sv: np.ndarray, observables: list, ops: Ops, trainable: Sequence[int],
                     apply_operations: bool = False) -> np.ndarray:
    """AdjointJacobianKokkos::adjointJacobian; returns jac[n_obs, n_tp]."""
    n = int(np.log2(sv.size))
    sv = np.ascontiguousarray(sv, dtype=np.complex128)
    hs = [_obs_handle(o) for o in observables]
    arr = (ctypes.c_void_p * max(1, len(hs)))(*[h.ptr.value for h in hs])
    oh = _OpsHandle(ops)
    tp = _as_i64(trainable)
    jac = np.zeros(len(hs) * len(tp), dtype=np.float64)
    _check(lib().orc_adjoint_jacobian(_dp(sv), n, arr, len(hs), oh.ptr, _ip(tp), len(tp),
                                      int(bool(apply_operations)), _dp(jac)))
    return jac.reshape(len(hs), len(tp))
