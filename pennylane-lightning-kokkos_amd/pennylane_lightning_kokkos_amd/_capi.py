"""ctypes front end of the C ABI declared in include/plk_amd.h.

This is the thin Python view of the drop-in boundary that tests and tools
call directly; the pybind11 module ``lightning_kokkos_qubit_ops`` exposes the
reference's class names on top of the same C functions. Loading fails loudly
when the HIP library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libplk_amd.so")

PLK_OK = 0
PLK_C64 = 64
PLK_C128 = 128

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_size_p = ctypes.POINTER(ctypes.c_size_t)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)
_c_int32_p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p

# name -> (restype, argtypes) for every entry point of include/plk_amd.h
SIGNATURES = {
    "plk_last_error": (ctypes.c_char_p, []),
    "plk_print_configuration": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "plk_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "plk_state_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "plk_state_destroy": (ctypes.c_int, [_vp]),
    "plk_state_num_qubits": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "plk_state_length": (ctypes.c_int, [_vp, _c_size_p]),
    "plk_state_dtype": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "plk_state_device_ptr": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "plk_state_upload": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "plk_state_download": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "plk_state_reset": (ctypes.c_int, [_vp]),
    "plk_state_set_basis": (ctypes.c_int, [_vp, ctypes.c_size_t]),
    "plk_state_set_values": (ctypes.c_int, [_vp, _c_size_p, _vp, ctypes.c_size_t]),
    "plk_state_copy": (ctypes.c_int, [_vp, _vp]),
    "plk_state_synchronize": (ctypes.c_int, [_vp]),
    "plk_apply": (ctypes.c_int, [_vp, ctypes.c_char_p, _c_size_p, ctypes.c_size_t, ctypes.c_int,
                                 _c_double_p, ctypes.c_size_t]),
    "plk_apply_matrix": (ctypes.c_int, [_vp, _c_double_p, _c_size_p, ctypes.c_size_t, ctypes.c_int]),
    "plk_apply_ops": (ctypes.c_int, [_vp, _vp]),
    "plk_apply_ops_blocks": (ctypes.c_int, [_vp, _vp, _c_size_p]),
    "plk_apply_generator": (ctypes.c_int, [_vp, ctypes.c_char_p, _c_size_p, ctypes.c_size_t,
                                           ctypes.c_int, _c_double_p]),
    "plk_expval_named": (ctypes.c_int, [_vp, ctypes.c_char_p, _c_size_p, ctypes.c_size_t, _c_double_p]),
    "plk_expval_matrix": (ctypes.c_int, [_vp, _c_double_p, _c_size_p, ctypes.c_size_t, _c_double_p]),
    "plk_expval_csr": (ctypes.c_int, [_vp, _c_double_p, _c_int64_p, _c_int64_p, ctypes.c_size_t,
                                      ctypes.c_size_t, _c_double_p]),
    "plk_expval_csr_i32": (ctypes.c_int, [_vp, _c_double_p, _c_int32_p, _c_int32_p, ctypes.c_size_t,
                                          ctypes.c_size_t, _c_double_p]),
    "plk_expval_obs": (ctypes.c_int, [_vp, _vp, _c_double_p]),
    "plk_expval_obs_noalloc": (ctypes.c_int, [_vp, _vp, _c_double_p]),
    "plk_inner_product": (ctypes.c_int, [_vp, _vp, _c_double_p, _c_double_p]),
    "plk_obs_named": (ctypes.c_int, [ctypes.c_char_p, _c_size_p, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "plk_obs_hermitian": (ctypes.c_int, [_c_double_p, _c_size_p, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "plk_obs_tensor": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "plk_obs_hamiltonian": (ctypes.c_int, [_c_double_p, ctypes.POINTER(_vp), ctypes.c_size_t,
                                           ctypes.POINTER(_vp)]),
    "plk_obs_sparse": (ctypes.c_int, [_c_double_p, _c_int64_p, _c_int64_p, ctypes.c_size_t,
                                      ctypes.c_size_t, _c_size_p, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "plk_obs_sparse_i32": (ctypes.c_int, [_c_double_p, _c_int32_p, _c_int32_p, ctypes.c_size_t,
                                          ctypes.c_size_t, _c_size_p, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "plk_obs_destroy": (ctypes.c_int, [_vp]),
    "plk_obs_wires": (ctypes.c_int, [_vp, _c_size_p, ctypes.c_size_t, _c_size_p]),
    "plk_obs_apply": (ctypes.c_int, [_vp, _vp]),
    "plk_ops_create": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_char_p), _c_double_p,
                                      _c_size_p, _c_size_p, _c_size_p, ctypes.POINTER(ctypes.c_int),
                                      _c_double_p, _c_size_p, ctypes.POINTER(_vp)]),
    "plk_ops_destroy": (ctypes.c_int, [_vp]),
    "plk_ops_num_par_ops": (ctypes.c_int, [_vp, _c_size_p]),
    "plk_adjoint_jacobian": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_size_t, _vp, _c_size_p,
                                            ctypes.c_size_t, ctypes.c_int, _c_double_p]),
    "plk_adjoint_batched": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_vp),
                                           ctypes.POINTER(_vp), ctypes.POINTER(_c_size_p),
                                           _c_size_p, _c_double_p, _c_double_p]),
    "plk_probs": (ctypes.c_int, [_vp, _c_size_p, ctypes.c_size_t, _c_double_p]),
    "plk_generate_samples": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_uint64, _c_size_p]),
    "plk_profiling_enable": (ctypes.c_int, [ctypes.c_int]),
    "plk_profiling_get": (ctypes.c_int, [ctypes.c_char_p, _c_double_p, _c_double_p, _c_double_p]),
    "plk_profiling_names": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
}


class PLException(RuntimeError):
    """Error raised by the engine (the reference's PLException)."""


_lib = None


def load(path: Optional[str] = None):
    """Loads libplk_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is missing: build the HIP library first "
                          "(python pennylane-lightning-kokkos_amd/build.py)")
    L = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int):
    if rc != PLK_OK:
        raise PLException(load().plk_last_error().decode())


def _csr_indices(indices, indptr):
    """(indices, indptr, is_int32): int32 arrays stay int32, anything else
    becomes int64."""
    i = np.asarray(indices).reshape(-1)
    p = np.asarray(indptr).reshape(-1)
    i32 = i.dtype == np.int32 and p.dtype == np.int32
    t = np.int32 if i32 else np.int64
    return np.ascontiguousarray(i, dtype=t), np.ascontiguousarray(p, dtype=t), i32


def _sizes(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64).reshape(-1))


def _sp(a: np.ndarray):
    return a.ctypes.data_as(_c_size_p)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_c_double_p)


def _cplx_doubles(m) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(-1)).view(np.float64)


def print_configuration() -> str:
    buf = ctypes.create_string_buffer(8192)
    check(load().plk_print_configuration(buf, len(buf)))
    return buf.value.decode()


def device_count() -> int:
    c = ctypes.c_int()
    check(load().plk_device_count(ctypes.byref(c)))
    return c.value


class State:
    """A 2^n state vector resident in HBM."""

    def __init__(self, num_qubits: int, dtype=np.complex128, device_id: int = 0):
        self.dtype = np.dtype(dtype)
        code = PLK_C128 if self.dtype == np.complex128 else PLK_C64
        if self.dtype not in (np.complex64, np.complex128):
            raise ValueError(f"unsupported dtype {dtype}")
        h = _vp()
        check(load().plk_state_create(int(num_qubits), code, int(device_id), ctypes.byref(h)))
        self.h = h
        self.num_qubits = int(num_qubits)
        self.length = 1 << self.num_qubits

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().plk_state_destroy(self.h)
                self.h = None
        except Exception:  # pragma: no cover - interpreter shutdown
            pass

    def upload(self, host: np.ndarray):
        a = np.ascontiguousarray(host, dtype=self.dtype).reshape(-1)
        check(load().plk_state_upload(self.h, a.ctypes.data, a.size))

    def download(self) -> np.ndarray:
        out = np.empty(self.length, dtype=self.dtype)
        check(load().plk_state_download(self.h, out.ctypes.data, out.size))
        return out

    def reset(self):
        check(load().plk_state_reset(self.h))

    def set_basis(self, index: int):
        check(load().plk_state_set_basis(self.h, int(index)))

    def set_values(self, indices, values):
        idx = _sizes(indices)
        vals = np.ascontiguousarray(np.asarray(values, dtype=self.dtype).reshape(-1))
        check(load().plk_state_set_values(self.h, _sp(idx), vals.ctypes.data, idx.size))

    def copy_from(self, other: "State"):
        check(load().plk_state_copy(self.h, other.h))

    def synchronize(self):
        check(load().plk_state_synchronize(self.h))

    def probs(self, wires: Sequence[int] = ()) -> np.ndarray:
        """Marginal probabilities of `wires` (all wires when empty), in the
        reference's order (MeasuresKokkos::probs; see plk_probs)."""
        w = _sizes(wires)
        out = np.zeros(2 ** (w.size if w.size else self.num_qubits), dtype=np.float64)
        check(load().plk_probs(self.h, _sp(w), w.size, _dp(out)))
        return out

    def generate_samples(self, shots: int, seed: int = 5374857) -> np.ndarray:
        """shots x num_qubits array of 0/1 (wire 0 first)."""
        out = np.zeros(max(1, shots * self.num_qubits), dtype=np.uint64)
        check(load().plk_generate_samples(self.h, int(shots), int(seed) & (2**64 - 1),
                                          out.ctypes.data_as(_c_size_p)))
        return out[: shots * self.num_qubits].reshape(shots, self.num_qubits)

    def device_ptr(self) -> int:
        p = _vp()
        check(load().plk_state_device_ptr(self.h, ctypes.byref(p)))
        return p.value

    def apply(self, name: str, wires: Sequence[int], inverse: bool = False,
              params: Sequence[float] = ()):
        w = _sizes(wires)
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
        check(load().plk_apply(self.h, name.encode(), _sp(w), w.size, int(bool(inverse)), _dp(p), p.size))

    def apply_matrix(self, matrix, wires: Sequence[int], inverse: bool = False):
        w = _sizes(wires)
        m = _cplx_doubles(matrix)
        check(load().plk_apply_matrix(self.h, _dp(m), _sp(w), w.size, int(bool(inverse))))

    def apply_ops(self, ops: "Ops"):
        check(load().plk_apply_ops(self.h, ops.h))

    def apply_ops_blocks(self, ops: "Ops") -> int:
        """Apply ops fused into 5-qubit dense blocks on the matrix cores
        (complex64, >= 9 qubits); returns the number of block passes."""
        nb = ctypes.c_size_t()
        check(load().plk_apply_ops_blocks(self.h, ops.h, ctypes.byref(nb)))
        return nb.value

    def apply_generator(self, name: str, wires: Sequence[int], adjoint: bool = False) -> float:
        w = _sizes(wires)
        s = ctypes.c_double()
        check(load().plk_apply_generator(self.h, name.encode(), _sp(w), w.size, int(bool(adjoint)),
                                         ctypes.byref(s)))
        return s.value

    def expval_named(self, name: str, wires: Sequence[int]) -> float:
        w = _sizes(wires)
        out = ctypes.c_double()
        check(load().plk_expval_named(self.h, name.encode(), _sp(w), w.size, ctypes.byref(out)))
        return out.value

    def expval_matrix(self, matrix, wires: Sequence[int]) -> float:
        w = _sizes(wires)
        m = _cplx_doubles(matrix)
        out = ctypes.c_double()
        check(load().plk_expval_matrix(self.h, _dp(m), _sp(w), w.size, ctypes.byref(out)))
        return out.value

    def expval_csr(self, data, indices, indptr) -> float:
        """CSR expectation value; int32 index arrays (scipy's default, the
        complex64 serialisation) go through plk_expval_csr_i32 unwidened."""
        d = _cplx_doubles(data)
        i, p, i32 = _csr_indices(indices, indptr)
        out = ctypes.c_double()
        if i32:
            check(load().plk_expval_csr_i32(self.h, _dp(d), i.ctypes.data_as(_c_int32_p),
                                            p.ctypes.data_as(_c_int32_p), i.size, p.size - 1,
                                            ctypes.byref(out)))
        else:
            check(load().plk_expval_csr(self.h, _dp(d), i.ctypes.data_as(_c_int64_p),
                                        p.ctypes.data_as(_c_int64_p), i.size, p.size - 1,
                                        ctypes.byref(out)))
        return out.value

    def expval(self, ob: "Obs") -> float:
        out = ctypes.c_double()
        check(load().plk_expval_obs(self.h, ob.h, ctypes.byref(out)))
        return out.value

    def expval_noalloc(self, ob: "Obs") -> float:
        """plk_expval_obs_noalloc: Pauli sums only, no device allocation."""
        out = ctypes.c_double()
        check(load().plk_expval_obs_noalloc(self.h, ob.h, ctypes.byref(out)))
        return out.value

    def inner(self, other: "State") -> complex:
        re = ctypes.c_double()
        im = ctypes.c_double()
        check(load().plk_inner_product(self.h, other.h, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def apply_obs(self, ob: "Obs"):
        check(load().plk_obs_apply(self.h, ob.h))


class Obs:
    """An observable handle (Named, Hermitian, Tensor, Hamiltonian, Sparse)."""

    def __init__(self, handle, keep=()):
        self.h = handle
        self._keep = list(keep)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().plk_obs_destroy(self.h)
                self.h = None
        except Exception:  # pragma: no cover
            pass

    @staticmethod
    def named(name: str, wires: Sequence[int]) -> "Obs":
        w = _sizes(wires)
        h = _vp()
        check(load().plk_obs_named(name.encode(), _sp(w), w.size, ctypes.byref(h)))
        return Obs(h)

    @staticmethod
    def hermitian(matrix, wires: Sequence[int]) -> "Obs":
        w = _sizes(wires)
        m = _cplx_doubles(matrix)
        h = _vp()
        check(load().plk_obs_hermitian(_dp(m), _sp(w), w.size, ctypes.byref(h)))
        return Obs(h)

    @staticmethod
    def tensor(factors: List["Obs"]) -> "Obs":
        arr = (_vp * max(1, len(factors)))(*[f.h.value for f in factors])
        h = _vp()
        check(load().plk_obs_tensor(arr, len(factors), ctypes.byref(h)))
        return Obs(h)

    @staticmethod
    def hamiltonian(coeffs, terms: List["Obs"]) -> "Obs":
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64).reshape(-1))
        arr = (_vp * max(1, len(terms)))(*[t.h.value for t in terms])
        h = _vp()
        check(load().plk_obs_hamiltonian(_dp(c), arr, len(terms), ctypes.byref(h)))
        return Obs(h)

    @staticmethod
    def sparse(data, indices, indptr, wires: Sequence[int]) -> "Obs":
        d = _cplx_doubles(data)
        i, p, i32 = _csr_indices(indices, indptr)
        w = _sizes(wires)
        h = _vp()
        if i32:
            check(load().plk_obs_sparse_i32(_dp(d), i.ctypes.data_as(_c_int32_p),
                                            p.ctypes.data_as(_c_int32_p), i.size, p.size - 1,
                                            _sp(w), w.size, ctypes.byref(h)))
        else:
            check(load().plk_obs_sparse(_dp(d), i.ctypes.data_as(_c_int64_p),
                                        p.ctypes.data_as(_c_int64_p), i.size, p.size - 1,
                                        _sp(w), w.size, ctypes.byref(h)))
        return Obs(h)

    def wires(self) -> List[int]:
        buf = np.zeros(256, dtype=np.uint64)
        cnt = ctypes.c_size_t()
        check(load().plk_obs_wires(self.h, _sp(buf), buf.size, ctypes.byref(cnt)))
        return [int(x) for x in buf[: cnt.value]]


class Ops:
    """A serialized operation list (the reference's OpsData)."""

    def __init__(self, names, params, wires, inverses, matrices=None):
        n = len(names)
        matrices = matrices if matrices is not None else [[] for _ in range(n)]
        if not (len(params) == len(wires) == len(inverses) == len(matrices) == n):
            raise ValueError("operation lists must have equal lengths")
        poff = np.zeros(n + 1, dtype=np.uint64)
        woff = np.zeros(n + 1, dtype=np.uint64)
        moff = np.zeros(n + 1, dtype=np.uint64)
        for i in range(n):
            poff[i + 1] = poff[i] + np.asarray(params[i]).size
            woff[i + 1] = woff[i] + len(wires[i])
            moff[i + 1] = moff[i] + np.asarray(matrices[i]).size
        pflat = np.ascontiguousarray(np.concatenate(
            [np.asarray(p, dtype=np.float64).reshape(-1) for p in params] + [np.zeros(1)]))
        wflat = np.ascontiguousarray(np.concatenate(
            [np.asarray(w, dtype=np.uint64).reshape(-1) for w in wires] + [np.zeros(1, np.uint64)]))
        mflat = np.ascontiguousarray(np.concatenate(
            [np.asarray(m, dtype=np.complex128).reshape(-1) for m in matrices]
            + [np.zeros(1, np.complex128)])).view(np.float64)
        cnames = (ctypes.c_char_p * max(1, n))(*[s.encode() for s in names])
        inv = (ctypes.c_int * max(1, n))(*[int(bool(b)) for b in inverses])
        h = _vp()
        check(load().plk_ops_create(n, cnames, _dp(pflat), _sp(poff), _sp(wflat), _sp(woff), inv,
                                    _dp(mflat), _sp(moff), ctypes.byref(h)))
        self.h = h
        self.names = list(names)
        self._num_par = int(np.count_nonzero(np.diff(poff)))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().plk_ops_destroy(self.h)
                self.h = None
        except Exception:  # pragma: no cover
            pass

    def num_par_ops(self, query: bool = False) -> int:
        """getNumParOps (ops with a non-empty parameter list); counted at
        creation unless query asks the library."""
        if not query:
            return self._num_par
        c = ctypes.c_size_t()
        check(load().plk_ops_num_par_ops(self.h, ctypes.byref(c)))
        return c.value


def adjoint_jacobian(state: State, observables: List[Obs], ops: Ops, trainable: Sequence[int],
                     apply_operations: bool = False) -> np.ndarray:
    tp = _sizes(trainable)
    nobs = len(observables)
    arr = (_vp * max(1, nobs))(*[o.h.value for o in observables])
    jac = np.zeros(max(1, nobs * tp.size), dtype=np.float64)
    check(load().plk_adjoint_jacobian(state.h, arr, nobs, ops.h, _sp(tp), tp.size,
                                      int(bool(apply_operations)), _dp(jac)))
    return jac[: nobs * tp.size].reshape(nobs, tp.size)


def adjoint_batched(num_qubits: Sequence[int], ops: Sequence[Ops], observables: Sequence[Obs],
                    trainable: Sequence[Sequence[int]] = None, dtype=np.complex64,
                    device_id: int = 0):
    """Expectation values and Jacobian rows of independent circuits run from
    |0...0> in one batched launch. Returns (expvals, [row per circuit])."""
    nc = len(num_qubits)
    if len(ops) != nc or len(observables) != nc:
        raise ValueError("num_qubits, ops and observables must have the same length")
    code = PLK_C128 if np.dtype(dtype) == np.complex128 else PLK_C64
    nq = (ctypes.c_int * max(1, nc))(*[int(q) for q in num_qubits])
    oarr = (_vp * max(1, nc))(*[o.h.value for o in ops])
    barr = (_vp * max(1, nc))(*[b.h.value for b in observables])
    if trainable is None:
        lens = [o.num_par_ops() for o in ops]
        tarr = None
        ntr = None
        keep = []
    else:
        keep = [_sizes(t) for t in trainable]
        lens = [k.size for k in keep]
        tarr = (_c_size_p * max(1, nc))(*[_sp(k) for k in keep])
        ntr = _sizes(lens)
    total = int(sum(lens))
    expvals = np.zeros(max(1, nc), dtype=np.float64)
    jac = np.zeros(max(1, total), dtype=np.float64)
    check(load().plk_adjoint_batched(code, int(device_id), nc, nq, oarr, barr, tarr,
                                     None if ntr is None else _sp(ntr), _dp(expvals), _dp(jac)))
    offs = np.concatenate(([0], np.cumsum(lens, dtype=np.int64))).tolist()
    rows = [jac[a:b] for a, b in zip(offs[:-1], offs[1:])]
    return expvals[:nc], rows


def profiling_enable(on: bool = True):
    check(load().plk_profiling_enable(int(bool(on))))


def profiling_get(kernel: str):
    a = ctypes.c_double()
    b = ctypes.c_double()
    c = ctypes.c_double()
    check(load().plk_profiling_get(kernel.encode(), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return {"launches": a.value, "total_ms": b.value, "bytes": c.value}


def profiling_names() -> List[str]:
    buf = ctypes.create_string_buffer(4096)
    check(load().plk_profiling_names(buf, len(buf)))
    s = buf.value.decode()
    return [x for x in s.split(",") if x]
