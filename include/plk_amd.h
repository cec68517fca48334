/*
 * plk_amd.h — C ABI of the MI355X (gfx950) state-vector engine.
 *
 * This is the drop-in boundary for the adjoint-Jacobian path of
 * PennyLane-Lightning-Kokkos. In the reference the boundary is the pybind11
 * module `lightning_kokkos_qubit_ops`
 * (pennylane_lightning_kokkos/src/bindings/Bindings.cpp:830-968); every entry
 * point below names the reference member it replaces. The C++ host layer of
 * this project (its own pybind11 module of the same name) and any other FFI
 * (ctypes, cgo, JNI) call these functions.
 *
 * Conventions
 *   - Every function returns PLK_OK (0) or an error code; the message of the
 *     last error on the calling thread is returned by plk_last_error(). The
 *     messages of errors the reference raises through PL_ABORT carry the
 *     reference's text (util/Error.hpp:27-142).
 *   - Wire w is bit (n-1-w) of the amplitude index (GateFunctors.hpp:32); for
 *     a k-wire operation wires[0] is the most significant bit of the local
 *     index, so dense matrices are row-major with index sum_p bit(wires[p])
 *     << (k-1-p) (GateFunctors.hpp:175-189).
 *   - Gate parameters, matrices, coefficients and all results are FP64 at the
 *     boundary, whatever the state's precision. Matrices are interleaved
 *     (re, im) doubles, row-major.
 *   - Host state buffers passed to upload/download are interleaved complex in
 *     the state's own precision (complex64 or complex128).
 *   - Work runs on one HIP stream per device owned by the library. Calls that
 *     return values synchronise that stream; gate applications are enqueued
 *     and may return before the kernels finish.
 */
#ifndef PLK_AMD_H
#define PLK_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PLK_OK = 0,
    PLK_ERR_INVALID = 1,     /* bad argument; reference PL_ABORT cases */
    PLK_ERR_UNSUPPORTED = 2, /* operation / observable not supported */
    PLK_ERR_HIP = 3,         /* HIP runtime error (launch, copy) */
    PLK_ERR_OOM = 4          /* device allocation failed */
} plk_status;

typedef enum { PLK_C64 = 64, PLK_C128 = 128 } plk_dtype;

typedef struct plk_state plk_state; /* a 2^n state vector in HBM */
typedef struct plk_obs plk_obs;     /* an observable tree */
typedef struct plk_ops plk_ops;     /* a serialized operation list */

/* ---- library ---------------------------------------------------------- */

/* Thread-local message of the last failing call ("" if none). */
const char *plk_last_error(void);

/* Writes a human-readable description of the device, architecture, build
 * flags and kernel configuration into buf (NUL-terminated, truncated to len).
 * Replaces print_configuration / kokkos_config_info (Bindings.cpp:842-852). */
int plk_print_configuration(char *buf, size_t len);

/* Number of visible HIP devices. */
int plk_device_count(int *out);

/* ---- state vector: LightningKokkos_C{64,128} (Bindings.cpp:59-585) ----- */

/* Allocates 2^num_qubits amplitudes of the given dtype on device_id and sets
 * |0...0>. Replaces the (num_qubits[, InitializationSettings]) constructors
 * (Bindings.cpp:63-70; StateVectorKokkos.hpp:140-478). */
int plk_state_create(int num_qubits, int dtype, int device_id,
                     plk_state **out);
int plk_state_destroy(plk_state *sv);
int plk_state_num_qubits(const plk_state *sv, int *out); /* numQubits */
int plk_state_length(const plk_state *sv, size_t *out);  /* dataLength */
int plk_state_dtype(const plk_state *sv, int *out);
/* Device address of the amplitudes (for zero-copy interop). */
int plk_state_device_ptr(plk_state *sv, void **out);

/* HostToDevice (Bindings.cpp:566-582; StateVectorKokkos.hpp:1618-1622).
 * length must equal dataLength. */
int plk_state_upload(plk_state *sv, const void *host, size_t length);
/* DeviceToHost (Bindings.cpp:554-565; StateVectorKokkos.hpp:1624-1628). */
int plk_state_download(const plk_state *sv, void *host, size_t length);
/* resetKokkos: |0...0> (Bindings.cpp:585; StateVectorKokkos.hpp:528-532). */
int plk_state_reset(plk_state *sv);
/* setBasisState (Bindings.cpp:86-91; StateVectorKokkos.hpp:492-495). */
int plk_state_set_basis(plk_state *sv, size_t index);
/* setStateVector: zero the state, then scatter values (complex, the state's
 * precision) at indices (Bindings.cpp:92-108; StateVectorKokkos.hpp:503-521). */
int plk_state_set_values(plk_state *sv, const size_t *indices,
                         const void *values, size_t count);
/* DeviceToDevice / updateData (StateVectorKokkos.hpp:1596-1616). */
int plk_state_copy(plk_state *dst, const plk_state *src);
/* Blocks until all work enqueued on the state's device has finished. */
int plk_state_synchronize(const plk_state *sv);

/* ---- gates: named methods, apply, applyGenerator (Bindings.cpp:109-432) - */

/* Apply one named gate (PauliX ... DoubleExcitationPlus, Rot, CRot, MultiRZ,
 * Identity). Replaces StateVectorKokkos::applyOperation(name, wires, inverse,
 * params) (StateVectorKokkos.hpp:585-600) and the per-gate bound methods. */
int plk_apply(plk_state *sv, const char *name, const size_t *wires,
              size_t num_wires, int inverse, const double *params,
              size_t num_params);
/* Dense 2^k x 2^k matrix on k wires; inverse applies the conjugate
 * transpose. Replaces apply(name, wires, inv, params, matrix)
 * (Bindings.cpp:243-261; StateVectorKokkos.hpp:704-796). */
int plk_apply_matrix(plk_state *sv, const double *matrix, const size_t *wires,
                     size_t num_wires, int inverse);
/* Apply every operation of ops in order, fusing runs of gates into single
 * HBM passes. Replaces apply(names, wires, adjoints, params)
 * (Bindings.cpp:233-242; StateVectorKokkos.hpp:640-676). */
int plk_apply_ops(plk_state *sv, const plk_ops *ops);
/* plk_apply_ops on a state of at least 9 qubits with runs of operations
 * fused into dense 32x32 blocks on five wires, each block one pass of a
 * matrix-core (MFMA) kernel: FP32 MFMA for complex64 states, FP64 MFMA for
 * complex128 states. Replaces the per-gate launches of
 * apply(names, wires, adjoints, params) (StateVectorKokkos.hpp:640-676) and
 * the dense multiQubitOpFunctor path (GateFunctors.hpp:195-300) for
 * circuits of one-to-five-wire gates. *num_blocks (may be NULL) receives
 * the number of block passes. */
int plk_apply_ops_blocks(plk_state *sv, const plk_ops *ops, size_t *num_blocks);
/* Apply the generator G of a parametric gate in place and return the scale s
 * with U(theta) = exp(i s theta G). The adjoint flag is accepted and ignored,
 * as in the reference. Replaces applyGenerator (Bindings.cpp:262-265;
 * StateVectorKokkos.hpp:687-695, 1275-1580). */
int plk_apply_generator(plk_state *sv, const char *name, const size_t *wires,
                        size_t num_wires, int adjoint, double *scale);

/* ---- measurements: ExpectationValue (Bindings.cpp:434-516) -------------- */

/* Named observable (Identity, PauliX/Y/Z, Hadamard) on wires
 * (MeasuresKokkos.hpp:80-101, 167-271). */
int plk_expval_named(const plk_state *sv, const char *name,
                     const size_t *wires, size_t num_wires, double *out);
/* Dense Hermitian matrix on wires (MeasuresKokkos.hpp:112-121, 283-346). */
int plk_expval_matrix(const plk_state *sv, const double *matrix,
                      const size_t *wires, size_t num_wires, double *out);
/* CSR matrix over the full state: sum_r Re(conj(psi_r) sum_j A_rj psi_cj)
 * (MeasuresKokkos.hpp:132-157; ExpValFunctors.hpp:252-280). data holds nnz
 * complex values (interleaved FP64); indptr holds num_rows + 1 entries. */
int plk_expval_csr(const plk_state *sv, const double *data,
                   const int64_t *indices, const int64_t *indptr, size_t nnz,
                   size_t num_rows, double *out);
/* plk_expval_csr with 32-bit indices and indptr, as the reference's
 * serialiser sends them for complex64 (_serialize.py:136-154); the arrays
 * reach the kernels without widening. Malformed matrices (decreasing indptr,
 * an index outside [0, num_rows)) fail with PLK_ERR_INVALID, detected by the
 * kernels. */
int plk_expval_csr_i32(const plk_state *sv, const double *data,
                       const int32_t *indices, const int32_t *indptr,
                       size_t nnz, size_t num_rows, double *out);
/* Re<psi|O|psi> of an observable tree (MeasuresKokkos.hpp:354-360). */
int plk_expval_obs(const plk_state *sv, const plk_obs *ob, double *out);
/* plk_expval_obs for an observable that is a sum of at most 4096 products of
 * Identity / PauliX / PauliY / PauliZ / Hadamard, allocating no device
 * memory. No counterpart in the reference: it is for callers that cannot
 * call hipMalloc (HIP graph capture, tight serving loops, harnesses that
 * forbid allocation). The partial sums and term tables live in a fixed
 * 256 KiB device array that each GPU holds until the process ends; larger
 * sums fail with PLK_ERR_INVALID, other observables with
 * PLK_ERR_UNSUPPORTED. All calls on one GPU share that array: calls that
 * overlap on one GPU race (the library's own lock serialises its calls, and
 * each call waits for its kernels). Same kernels and result as
 * plk_expval_obs. */
int plk_expval_obs_noalloc(const plk_state *sv, const plk_obs *ob,
                           double *out);
/* <x|y> accumulated in FP64 (LinearAlgebraKokkos.hpp:155-236). */
int plk_inner_product(const plk_state *x, const plk_state *y, double *re,
                      double *im);

/* ---- observables (Bindings.cpp:591-736; ObservablesKokkos.hpp) ---------- */

/* NamedObsKokkos_C*(name, wires) (ObservablesKokkos.hpp:77-119). */
int plk_obs_named(const char *name, const size_t *wires, size_t num_wires,
                  plk_obs **out);
/* HermitianObsKokkos_C*(matrix, wires) (ObservablesKokkos.hpp:124-180). */
int plk_obs_hermitian(const double *matrix, const size_t *wires,
                      size_t num_wires, plk_obs **out);
/* TensorProdObsKokkos_C*(obs) (ObservablesKokkos.hpp:185-290); the factors'
 * wires must be disjoint. Factors are copied: the caller keeps its handles. */
int plk_obs_tensor(const plk_obs *const *factors, size_t num_factors,
                   plk_obs **out);
/* HamiltonianKokkos_C*(coeffs, obs) (ObservablesKokkos.hpp:296-402). */
int plk_obs_hamiltonian(const double *coeffs, const plk_obs *const *terms,
                        size_t num_terms, plk_obs **out);
/* SparseHamiltonianKokkos_C*(data, indices, indptr, wires)
 * (ObservablesKokkos.hpp:409-517). */
int plk_obs_sparse(const double *data, const int64_t *indices,
                   const int64_t *indptr, size_t nnz, size_t num_rows,
                   const size_t *wires, size_t num_wires, plk_obs **out);
/* SparseHamiltonianKokkos_C64 with the int32 indices of the complex64
 * serialisation (_serialize.py:136-154). Every sparse observable keeps one
 * device copy per device and state precision after its first use (values in
 * the state's precision, 32-bit indices whenever they fit). */
int plk_obs_sparse_i32(const double *data, const int32_t *indices,
                       const int32_t *indptr, size_t nnz, size_t num_rows,
                       const size_t *wires, size_t num_wires, plk_obs **out);
int plk_obs_destroy(plk_obs *ob);
/* Sorted wires of an observable (get_wires). *count receives the number of
 * wires; at most cap are written. */
int plk_obs_wires(const plk_obs *ob, size_t *wires, size_t cap,
                  size_t *count);
/* ObservableKokkos::applyInPlace (ObservablesKokkos.hpp:45). */
int plk_obs_apply(plk_state *sv, const plk_obs *ob);

/* ---- operations: OpsStructKokkos_C* / create_ops_list
 *      (Bindings.cpp:741-805; AdjointDiffKokkos.hpp:17-173) -------------- */

/* Flattened op list: op i has params[param_offsets[i] .. param_offsets[i+1]),
 * wires[wire_offsets[i] .. wire_offsets[i+1]) and
 * matrices[2*matrix_offsets[i] .. 2*matrix_offsets[i+1]) (complex entries;
 * empty unless the op has no named kernel). All offset arrays have
 * num_ops + 1 entries. */
int plk_ops_create(size_t num_ops, const char *const *names,
                   const double *params, const size_t *param_offsets,
                   const size_t *wires, const size_t *wire_offsets,
                   const int *inverses, const double *matrices,
                   const size_t *matrix_offsets, plk_ops **out);
int plk_ops_destroy(plk_ops *ops);
/* Number of operations with a non-empty parameter list (getNumParOps). */
int plk_ops_num_par_ops(const plk_ops *ops, size_t *out);

/* ---- adjoint Jacobian: AdjointJacobianKokkos_C*.adjoint_jacobian
 *      (Bindings.cpp:808-821; AdjointDiffKokkos.hpp:404-478) ------------- */

/* jac (num_obs x num_trainable, row-major, FP64) receives
 * d<O_j>/d theta_{trainable[t]} for the state sv (left unchanged). trainable
 * holds sorted indices into the parametric operations. If apply_operations is
 * non-zero the operations are first applied to a copy of sv. The backward
 * sweep recomputes the state in reverse with fused HIP kernels and reduces
 * every inner product in FP64. */
int plk_adjoint_jacobian(const plk_state *sv, const plk_obs *const *obs,
                         size_t num_obs, const plk_ops *ops,
                         const size_t *trainable, size_t num_trainable,
                         int apply_operations, double *jac);

/* Batched form of plk_adjoint_jacobian: num_circuits independent circuits,
 * each applied to |0...0> of num_qubits[c] (1..20) qubits, each with one
 * observable obs[c] (a sum of products of Identity / PauliX / PauliY /
 * PauliZ / Hadamard), all run by one persistent launch on device_id in the
 * precision dtype. Replaces a loop over circuits of
 * AdjointJacobianKokkos::adjointJacobian (AdjointDiffKokkos.hpp:404-478)
 * with apply_operations = true, the per-tape calls of the device's
 * adjoint_jacobian (lightning_kokkos.py:621-687). Ops are limited to gates
 * (and generators) on at most two wires.
 * expvals[c] receives <obs[c]>; the Jacobian rows are packed: circuit c's
 * row of num_trainable[c] entries starts after the rows of circuits 0..c-1.
 * trainable == NULL makes every parametric op of every circuit trainable
 * (row length = plk_ops_num_par_ops). Synchronous; FP64 results. */
int plk_adjoint_batched(int dtype, int device_id, size_t num_circuits,
                        const int *num_qubits, const plk_ops *const *ops,
                        const plk_obs *const *obs,
                        const size_t *const *trainable,
                        const size_t *num_trainable, double *expvals,
                        double *jac);

/* ---- probabilities and samples: LightningKokkos_C*.probs / GenerateSamples
 *      (Bindings.cpp:517-553; MeasuresKokkos.hpp:389-569) --------------- */

/* MeasuresKokkos::probs(wires) (MeasuresKokkos.hpp:413-520): the 2^num_wires
 * marginal probabilities of `wires`, FP64 sums, deterministic. For sorted
 * wires, output index bits follow the wires (wires[0] most significant). For
 * unsorted wires the reference's transposition is reproduced exactly: the
 * output is the marginal in the order wires[pi[pi[k]]], pi = argsort(wires)
 * (equal to the given order whenever that order is an involution).
 * num_wires == 0 means every wire in order (MeasuresKokkos::probs(),
 * 389-411). */
int plk_probs(const plk_state *sv, const size_t *wires, size_t num_wires,
              double *probs);
/* MeasuresKokkos::generate_samples (MeasuresKokkos.hpp:522-569): num_shots
 * samples of the full register by inverse transform of the FP64 exclusive
 * CDF. samples[s * n + (n - 1 - j)] = bit j of shot s's basis index (a row
 * per shot, wire 0 first), as the reference lays them out. Uniforms are
 * output s of SplitMix64 seeded with `seed` (the reference draws from
 * Kokkos' XorShift64 pool, so individual samples differ; distributions
 * agree). */
int plk_generate_samples(const plk_state *sv, size_t num_shots, uint64_t seed,
                         size_t *samples);

/* ---- measurement support ---------------------------------------------- */

/* Enables (1) or disables (0) per-kernel timing with HIP events on the
 * library's stream, and clears the accumulated statistics. */
int plk_profiling_enable(int on);
/* Accumulated statistics of one kernel family (see plk_profiling_names):
 * number of launches, total milliseconds and algorithmic HBM bytes moved. */
int plk_profiling_get(const char *kernel, double *launches, double *total_ms,
                      double *bytes);
/* Comma-separated names of the kernel families with statistics. */
int plk_profiling_names(char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* PLK_AMD_H */
