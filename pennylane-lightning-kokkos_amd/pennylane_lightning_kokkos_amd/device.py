"""A `lightning.kokkos`-style device over the gfx950 engine.

The reference device (pennylane_lightning_kokkos/lightning_kokkos.py:139-742)
is a PennyLane `QubitDevice`; PennyLane is not a dependency here, so this
class keeps the device's method names and semantics on plain objects:
operations are `torch_adjoint.Op` (name, wires, params, inverse, matrix),
observables are the tuple descriptions used across this package
(("named", name, wires), ("herm", matrix, wires), ("tensor", [...]),
("ham", coeffs, [...]), ("sparse", data, indices, indptr, wires)), and a
`Tape` holds operations, observables and trainable parameter indices.

Kept from the reference:
- serialization (`_serialize.py:252-308`): Rot expands to RZ(phi) RY(theta)
  RZ(omega); ops without a named kernel travel as their matrix (already
  inverted, inverse flag cleared); StatePrep / BasisState are not serialized
  and set `uses_stateprep`;
- `apply` (`lightning_kokkos.py:404-422`): a leading StatePrep / BasisState
  prepares the state, later ones raise;
- `adjoint_jacobian` (`lightning_kokkos.py:621-687`): expectation values
  only, no Hermitian observables, no multi-parameter ops but Rot; trainable
  parameters of StatePrep / BasisState get zero columns, the others are
  shifted by one when the tape starts with a state preparation;
- `vjp` (`lightning_kokkos.py:689-728`): dy^T J as one adjoint sweep of the
  Hamiltonian sum_j dy_j O_j;
- `batch_adjoint_jacobian`: the per-tape loop of a batch, with the tapes
  the batched engine takes run in one launch;
- `probability` rejects out-of-order wires as the device does
  (`tests/test_probs.py:116-191`); the engine itself supports them.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _capi
from .torch_adjoint import Op

STATE_PREPS = ("StatePrep", "QubitStateVector", "BasisState")


class QuantumFunctionError(Exception):
    """Raised for circuits the adjoint method cannot differentiate."""


class DeviceError(Exception):
    """Raised for operations the device cannot apply in this position."""


@dataclass
class Tape:
    operations: List[Op]
    observables: List[tuple]
    trainable_params: Optional[List[int]] = None

    def all_params(self):
        """(op index, parameter) for every parameter of every operation."""
        out = []
        for i, op in enumerate(self.operations):
            for p in op.params:
                out.append((i, p))
        return out

    def get_trainable(self) -> List[int]:
        if self.trainable_params is None:
            return list(range(len(self.all_params())))
        return sorted(self.trainable_params)


def serialize_ops(tape: Tape):
    """_serialize.py:252-308: (names, params, wires, inverses, matrices) and
    whether the tape prepares its state."""
    names, params, wires, inverses, mats = [], [], [], [], []
    uses_stateprep = False
    for o in tape.operations:
        if o.name in STATE_PREPS:
            uses_stateprep = True
            continue
        if o.name == "Rot" and not o.inverse:
            phi, theta, omega = o.params
            expanded = [Op("RZ", o.wires, [phi]), Op("RY", o.wires, [theta]),
                        Op("RZ", o.wires, [omega])]
        else:
            expanded = [o]
        for op in expanded:
            names.append(op.name)
            wires.append(list(op.wires))
            if op.matrix is not None:
                m = np.asarray(op.matrix, dtype=np.complex128)
                if op.inverse:
                    m = m.conj().T
                params.append([])
                mats.append(m.reshape(-1))
                inverses.append(False)
            else:
                params.append([float(p) for p in op.params])
                mats.append([])
                inverses.append(bool(op.inverse))
    return (names, params, wires, inverses, mats), uses_stateprep


def batch_eligible(serialized, uses_stateprep: bool, observables) -> bool:
    """Whether a serialized tape can run in the batched engine
    (`plk_adjoint_batched`): no state preparation, no explicit matrices,
    operations on at most two wires, exactly one observable that is a sum of
    Pauli words."""
    _, _, wires, _, mats = serialized
    return (not uses_stateprep and len(observables) == 1 and _pauli_only(observables[0]) and
            all(len(w) <= 2 for w in wires) and all(len(m) == 0 for m in mats))


_NAMED = {
    "Identity": np.eye(2, dtype=complex),
    "PauliX": np.array([[0, 1], [1, 0]], dtype=complex),
    "PauliY": np.array([[0, -1j], [1j, 0]], dtype=complex),
    "PauliZ": np.array([[1, 0], [0, -1]], dtype=complex),
    "Hadamard": np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2),
}


def _pauli_only(desc) -> bool:
    kind = desc[0]
    if kind == "named":
        return desc[1] in _NAMED
    if kind == "tensor":
        return all(_pauli_only(f) for f in desc[1])
    if kind == "ham":
        return all(_pauli_only(t) for t in desc[2])
    return False


def _obs_wires(desc) -> List:
    """Wires in order of first appearance (a Hamiltonian's wire order)."""
    kind = desc[0]
    if kind in ("named", "herm"):
        return list(desc[2])
    out = []
    for t in desc[1] if kind == "tensor" else desc[2]:
        for w in _obs_wires(t):
            if w not in out:
                out.append(w)
    return out


def _embed(U, w, W):
    """U acting on wires w, as a matrix on the wire list W (W[0] most
    significant)."""
    k, rest = len(W), [x for x in W if x not in w]
    full = np.kron(U, np.eye(2 ** len(rest), dtype=complex))
    order = list(w) + rest
    perm = [order.index(x) for x in W]
    t = full.reshape([2] * (2 * k)).transpose(perm + [k + p for p in perm])
    return t.reshape(2 ** k, 2 ** k)


def _obs_matrix(desc, W):
    """Dense matrix of an observable description on the wire list W."""
    kind = desc[0]
    if kind == "named":
        return _embed(_NAMED[desc[1]], list(desc[2]), W)
    if kind == "herm":
        return _embed(np.asarray(desc[1], dtype=complex), list(desc[2]), W)
    if kind == "tensor":
        M = np.eye(2 ** len(W), dtype=complex)
        for f in desc[1]:
            M = M @ _obs_matrix(f, W)
        return M
    if kind == "ham":
        return sum(c * _obs_matrix(t, W) for c, t in zip(desc[1], desc[2]))
    raise ValueError(f"no dense matrix for {kind}")


class LightningKokkosAMD:
    """lightning.kokkos-compatible device on one MI355X (or any gfx950)."""

    name = "Lightning Kokkos PennyLane plugin (gfx950)"
    short_name = "lightning.kokkos"

    def __init__(self, wires, *, c_dtype=np.complex128, shots=None, kokkos_args=None,
                 sync=True, batch_obs=False, device_id: int = 0):
        self.wires = list(range(wires)) if isinstance(wires, int) else list(wires)
        self.num_wires = len(self.wires)
        self.wire_map = {w: i for i, w in enumerate(self.wires)}
        self.C_DTYPE = np.dtype(c_dtype)
        if self.C_DTYPE not in (np.complex64, np.complex128):
            raise TypeError(f"Unsupported complex Type: {c_dtype}")
        self.use_csingle = self.C_DTYPE == np.complex64
        self.shots = shots
        self.kokkos_args = kokkos_args
        self._state_handle = _capi.State(self.num_wires, self.C_DTYPE, device_id)

    # ---- state ------------------------------------------------------------
    def reset(self):
        self._state_handle.reset()

    @property
    def state(self) -> np.ndarray:
        return self._state_handle.download()

    def syncH2D(self, state_vector):
        self._state_handle.upload(np.ravel(state_vector).astype(self.C_DTYPE))

    def syncD2H(self, state_vector):
        state_vector[:] = self._state_handle.download()

    def _map(self, wires) -> List[int]:
        return [self.wire_map[w] for w in wires]

    # ---- operations -------------------------------------------------------
    def apply(self, operations: Sequence[Op], **kwargs):
        ops = list(operations)
        if ops and ops[0].name in STATE_PREPS:
            self._prepare(ops[0])
            ops = ops[1:]
        for o in ops:
            if o.name in STATE_PREPS:
                raise DeviceError(
                    f"Operation {o.name} cannot be used after other Operations have already "
                    f"been applied on a {self.short_name} device.")
        ops = [o for o in ops if o.name != "Identity"]
        if not ops:
            return
        mapped = Tape([Op(o.name, self._map(o.wires), list(o.params), o.inverse, o.matrix)
                       for o in ops], [])
        (names, params, wires, inverses, mats), _ = serialize_ops(mapped)
        self._state_handle.apply_ops(_capi.Ops(names, params, wires, inverses, mats))

    def _prepare(self, op: Op):
        wires = self._map(op.wires)
        if op.name == "BasisState":
            bits = [int(b) for b in np.ravel(op.params[0])]
            index = 0
            for w, b in zip(wires, bits):
                index |= b << (self.num_wires - 1 - w)
            self._state_handle.set_basis(index)
            return
        vec = np.ravel(np.asarray(op.params[0], dtype=np.complex128))
        if len(wires) == self.num_wires and wires == sorted(wires):
            self.syncH2D(vec)
            return
        # A state on a subset of wires: the rest start in |0>.
        full = np.zeros(2**self.num_wires, dtype=np.complex128)
        for k, amp in enumerate(vec):
            idx = 0
            for pos, w in enumerate(wires):
                bit = (k >> (len(wires) - 1 - pos)) & 1
                idx |= bit << (self.num_wires - 1 - w)
            full[idx] = amp
        self.syncH2D(full)

    # ---- measurements -----------------------------------------------------
    def _obs(self, desc):
        kind = desc[0]
        if kind == "named":
            return _capi.Obs.named(desc[1], self._map(desc[2]))
        if kind == "herm":
            return _capi.Obs.hermitian(desc[1], self._map(desc[2]))
        if kind == "tensor":
            return _capi.Obs.tensor([self._obs(f) for f in desc[1]])
        if kind == "ham":
            return _capi.Obs.hamiltonian(desc[1], [self._obs(t) for t in desc[2]])
        if kind == "sparse":
            # Index width per precision as the reference serialises it
            # (_serialize.py:136-154): int32 for complex64, int64 for complex128.
            rtype = np.int32 if self.use_csingle else np.int64
            return _capi.Obs.sparse(desc[1], np.asarray(desc[2]).astype(rtype),
                                    np.asarray(desc[3]).astype(rtype), self._map(desc[4]))
        raise ValueError(f"unknown observable {kind}")

    def expval(self, observable) -> float:
        """Expectation value (lightning_kokkos.py:507-559). A Hamiltonian on
        fewer than 13 wires that is not a sum of Pauli words goes through
        its dense matrix, as the reference routes every such Hamiltonian
        (521-534); a sum of Pauli words keeps the Pauli-word kernels, whose
        value is the same and whose cost is linear in the terms."""
        if observable[0] == "ham" and not _pauli_only(observable):
            wires = _obs_wires(observable)
            if len(wires) < 13:
                return self._state_handle.expval_matrix(_obs_matrix(observable, wires),
                                                        self._map(wires))
        return self._state_handle.expval(self._obs(observable))

    def probability(self, wires=None):
        wires = self.wires if wires is None else list(wires)
        mapped = self._map(wires)
        if mapped != sorted(mapped):
            raise RuntimeError(
                "Lightning does not currently support out-of-order indices for probabilities")
        return self._state_handle.probs(mapped)

    def generate_samples(self, seed: int = 5374857) -> np.ndarray:
        if self.shots is None:
            raise DeviceError("generate_samples needs a finite number of shots")
        return self._state_handle.generate_samples(int(self.shots), seed).astype(np.int64)

    def execute(self, tape: Tape) -> np.ndarray:
        self.apply(tape.operations)
        return np.array([self.expval(o) for o in tape.observables])

    # ---- differentiation --------------------------------------------------
    @staticmethod
    def _check_observables(observables):
        def walk(d):
            if d[0] == "herm":
                raise QuantumFunctionError(
                    "LightningKokkos adjoint differentiation method does not currently support "
                    "the Hermitian observable")
            if d[0] in ("tensor", "ham"):
                for t in d[1 if d[0] == "tensor" else 2]:
                    walk(t)

        for d in observables:
            walk(d)

    @staticmethod
    def _check_operations(operations):
        for op in operations:
            if len(op.params) > 1 and op.name != "Rot":
                raise QuantumFunctionError(
                    f"The {op.name} operation is not supported using the \"adjoint\" "
                    "differentiation method")

    def adjoint_jacobian(self, tape: Tape, starting_state=None, use_device_state=False):
        """d<O_j>/d(trainable parameter), shape (n_obs, n_trainable), with the
        reference's column mapping (zero columns for state-preparation
        parameters)."""
        self._check_observables(tape.observables)
        trainable = tape.get_trainable()
        if not trainable:
            return np.array(0)
        self._check_operations(tape.operations)
        if starting_state is not None:
            self.syncH2D(starting_state)
        elif not use_device_state:
            self.reset()
            self.execute(tape)
        mapped_ops = [Op(o.name, self._map(o.wires), list(o.params), o.inverse, o.matrix)
                      for o in tape.operations]
        ser, uses_sp = serialize_ops(Tape(mapped_ops, []))
        owners = tape.all_params()
        tp_shift, record_rows = [], []
        for col, tp in enumerate(trainable):
            op = tape.operations[owners[tp][0]]
            if op.name not in STATE_PREPS:
                tp_shift.append(tp)
                record_rows.append(col)
        if uses_sp:
            tp_shift = [t - 1 for t in tp_shift]
        obs = [self._obs(d) for d in tape.observables]
        jac = np.zeros((len(obs), len(trainable)))
        if tp_shift:
            part = _capi.adjoint_jacobian(self._state_handle, obs, _capi.Ops(*ser), tp_shift,
                                          False)
            jac[:, record_rows] = part
        return jac

    def batch_adjoint_jacobian(self, tapes: Sequence[Tape]) -> List[np.ndarray]:
        """Jacobians of many tapes, each run from |0...0> (the reference device
        loops `adjoint_jacobian` over the tapes of a batch,
        lightning_kokkos.py:621-687 per tape). Tapes the batched engine takes
        -- no state preparation, no explicit matrices, operations on at most
        two wires, one observable that is a sum of Pauli words -- run together
        in one `plk_adjoint_batched` launch in the device's precision; the
        others go through `adjoint_jacobian`. Returns one array per tape, as
        `adjoint_jacobian` returns it."""
        out: List[Optional[np.ndarray]] = [None] * len(tapes)
        batch_ops, batch_obs, batch_tp, where = [], [], [], []
        for k, tape in enumerate(tapes):
            self._check_observables(tape.observables)
            trainable = tape.get_trainable()
            if not trainable:
                out[k] = np.array(0)
                continue
            self._check_operations(tape.operations)
            mapped = [Op(o.name, self._map(o.wires), list(o.params), o.inverse, o.matrix)
                      for o in tape.operations]
            ser, uses_sp = serialize_ops(Tape(mapped, []))
            names, params, wires, inverses, mats = ser
            if not batch_eligible(ser, uses_sp, tape.observables):
                out[k] = self.adjoint_jacobian(tape)
                continue
            batch_ops.append(_capi.Ops(names, params, wires, inverses, mats))
            batch_obs.append(self._obs(tape.observables[0]))
            batch_tp.append(trainable)
            where.append(k)
        if where:
            _, rows = _capi.adjoint_batched([self.num_wires] * len(where), batch_ops, batch_obs,
                                            batch_tp, dtype=self.C_DTYPE)
            for k, row in zip(where, rows):
                out[k] = np.asarray(row, dtype=np.float64).reshape(1, -1)
        return out

    def vjp(self, observables, dy, starting_state=None, use_device_state=False):
        """Processing function tape -> dy^T J (one sweep of sum_j dy_j O_j)."""
        dy = np.asarray(dy)
        if np.allclose(dy, 0):
            return lambda tape: np.zeros(len(tape.get_trainable()))
        if len(dy) != len(observables):
            raise ValueError("Number of observables in the tape must be the same as the length "
                             "of dy in the vjp method")
        if np.iscomplexobj(dy):
            raise ValueError("The vjp method only works with a real-valued dy when the tape is "
                             "returning an expectation value")
        ham = ("ham", [float(x) for x in dy], list(observables))

        def processing_fn(tape: Tape):
            if not tape.get_trainable():
                return np.array([], dtype=np.float64)
            new = Tape(tape.operations, [ham], tape.trainable_params)
            return self.adjoint_jacobian(new, starting_state, use_device_state)[0]

        return processing_fn
