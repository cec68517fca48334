// Internal declarations of the gfx950 state-vector engine (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "plk_amd.h"

namespace plk {

using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// Errors: every C entry point converts an Error into a status code plus a
// thread-local message (the reference's PL_ABORT -> LightningException path,
// util/Error.hpp:27-142).
// ---------------------------------------------------------------------------
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void raise(int code, const std::string &msg);
void check_hip(hipError_t e, const char *what);

#define PLK_HIP(expr) ::plk::check_hip((expr), #expr)

// ---------------------------------------------------------------------------
// Device context: one stream, scratch for reductions, kernel statistics.
// ---------------------------------------------------------------------------
struct KernelStats {
    double launches = 0;
    double ms = 0;
    double bytes = 0;
};

struct Device {
    int id = 0;
    hipStream_t stream = nullptr;
    double *red = nullptr;        // per-block FP64 partials
    size_t red_cap = 0;           // doubles
    double *red_out = nullptr;    // final reduction results
    size_t red_out_cap = 0;
    int num_cus = 256;
    void *work = nullptr;         // grow-only device workspace
    size_t work_cap = 0;          // bytes
    double *jac = nullptr;        // grow-only Jacobian accumulator
    size_t jac_cap = 0;           // doubles
    void *noalloc_scratch = nullptr; // this device's copy of the fixed scratch array
    // Per-call arena for uploaded plan tables: bump allocation, rewound at
    // the first allocation of each entry-point call (stream order keeps a
    // rewound region from being overwritten before earlier kernels read it).
    void *arena = nullptr;
    size_t arena_cap = 0;
    size_t arena_used = 0;
    uint64_t arena_call = 0;
    std::vector<void *> arena_retired; // outgrown chunks, freed at the next rewind
};

// Marks the start of an entry-point call (rewinds every device's arena
// lazily at its next allocation).
void begin_call();
// size bytes of the current call's arena on d (256-byte aligned).
void *arena_alloc(Device &d, size_t bytes);

Device &device(int id);
// Serializes every entry point of the C ABI: the per-device scratch buffers
// (partials, workspace, Jacobian), the temporary-state pool and the launch
// caches are shared by all callers of the library.
std::recursive_mutex &engine_mutex();
// hipMalloc on the current device; on failure the device's pool of recycled
// temporary states is released and the allocation retried once.
void *dev_malloc(int device, size_t bytes, const char *what);
void ensure_red(Device &d, size_t partial_doubles, size_t out_doubles);
// Grows d.jac to at least `doubles` entries.
double *ensure_jac(Device &d, size_t doubles);
// Grows d.work to at least `bytes` (synchronizes the stream when it grows).
void *ensure_work(Device &d, size_t bytes);
void set_device(int id);

// Profiling: when enabled, launch() brackets kernels with HIP events.
bool profiling_on();
// Marks the device stream as being captured into a graph (profiling scopes
// record nothing meanwhile).
void set_capturing(bool on);
void record_kernel(const char *family, float ms, double bytes);

struct ProfScope {
    const char *family;
    double bytes;
    hipStream_t stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ProfScope(const char *f, double b, hipStream_t s);
    ~ProfScope();
};

// Host wall time of a scope, recorded as one launch of `family` (no bytes)
// while profiling is on: host planning, packing and copies next to the
// kernels they feed.
struct HostScope {
    const char *family;
    double t0 = -1.0;
    explicit HostScope(const char *f);
    ~HostScope();
};

// ---------------------------------------------------------------------------
// State vector.
// ---------------------------------------------------------------------------
struct State {
    int n = 0;
    int dtype = PLK_C128;
    size_t len = 0;
    void *d = nullptr;
    int device = 0;
    size_t elem_bytes() const { return dtype == PLK_C64 ? 8 : 16; }
    size_t bytes() const { return len * elem_bytes(); }
};

// RAII temporary device state of the same shape.
struct TmpState {
    State s;
    TmpState(int n, int dtype, int dev);
    ~TmpState();
    TmpState(const TmpState &) = delete;
    TmpState &operator=(const TmpState &) = delete;
};

// ---------------------------------------------------------------------------
// Compiled operations (host side, FP64).
// ---------------------------------------------------------------------------
enum class OpKind { Identity, Dense, Parity };

struct CompiledOp {
    std::string name;
    std::vector<size_t> wires;  // wires[0] = most significant local bit
    OpKind kind = OpKind::Dense;
    std::vector<cd> mat;        // Dense: 2^m x 2^m row-major
    cd even{1, 0}, odd{1, 0};   // Parity: phase by parity of the wire bits
};

struct CompiledGen {
    std::vector<size_t> wires;
    OpKind kind = OpKind::Dense; // Dense or Parity (signs +1 / -1)
    std::vector<cd> mat;
    double scale = 0;
};

bool is_named_gate(const std::string &name);
size_t named_gate_wires(const std::string &name); // 0 = any (MultiRZ)
size_t named_gate_params(const std::string &name);
bool has_generator(const std::string &name);

// U (or U^dagger when inverse) of a named gate or of a matrix op.
CompiledOp compile_op(const std::string &name, const std::vector<size_t> &wires,
                      bool inverse, const std::vector<double> &params,
                      const std::vector<cd> &matrix, int num_qubits);
CompiledGen compile_generator(const std::string &name,
                              const std::vector<size_t> &wires, int num_qubits);

// ---------------------------------------------------------------------------
// Kernel launchers (implemented in the .hip files).
// ---------------------------------------------------------------------------
constexpr int SEG_KMAX = 3;      // wires per register-fused segment
constexpr int SEG_PMAX = 4;      // trainable ops per segment
constexpr int SEG_NJMAX = 8;     // observable states per launch
constexpr int SEG_BLOCK = 256;

// Structure of an operation's matrix, chosen on the host so the kernels skip
// the multiplications a dense complex product would waste.
enum SegUKind : int32_t { U_DENSE = 0, U_DIAG = 1, U_REAL = 2, U_SWAP = 3 };
// Structure of a generator: dense, real diagonal, (one wire) X or Y, or the
// parity sign (-1)^popcount of the op-local index (Z, Z(x)Z).
enum SegGKind : int32_t { G_DENSE = 0, G_DIAG = 1, G_X = 2, G_Y = 3, G_Z = 4 };

// Host-side description of one fused operation: `cls` selects the
// compile-time local-position case, m is 1 or 2 wires, U and G are row-major
// complex matrices (4x4, or 2x2 in the first 8 reals). Packed for the device
// on the host into an op group (see OpHead).
template <typename T> struct SegOp {
    int32_t cls;
    int32_t slot;  // trainable slot in [0, SEG_PMAX) or -1
    int32_t ukind; // SegUKind
    int32_t gkind; // SegGKind
    int32_t swap;  // U_SWAP: 4 * a + b, local indices a < b exchanged
    int32_t m;     // wires of the op
    int32_t pad[2];
    T U[32];
    T G[32];
};

// Device op group (a segment or a bundle): nops 16-byte headers, then the
// payloads. A header's offsets are bytes from the group start; the payloads
// hold only what the kind reads:
//   U_DENSE 2 dim^2 reals (row-major complex), U_DIAG 2 dim (complex
//   diagonal), U_REAL dim^2 (row-major real), U_SWAP none;
//   G_DENSE 2 dim^2, G_DIAG dim (real diagonal), G_X / G_Y / G_Z none.
struct OpHead {
    int32_t cls;
    int32_t slot;
    int32_t kinds; // ukind | gkind << 8 | swap << 16
    int32_t offs;  // U offset | G offset << 16
};

struct SegLaunch {
    int n;              // qubits
    int K;              // segment wires
    int bits[SEG_KMAX]; // ascending bit positions of the segment wires
    int nops;
    int nslots;
    int write_phi;
    int nj;             // observable states in this launch
};

// Forward/backward fused segment over phi and nj observable states. Writes
// per-block FP64 partials [slot][j][block] into dev.red when nj > 0 and
// nslots > 0; returns the grid size used.
int launch_segment(Device &dev, const State &phi, void *const *h_ptrs,
                   const void *d_ops, const SegLaunch &L);
// Largest nj accepted by launch_segment for a dtype and segment width.
int seg_max_nj(int dtype, int K);
// Reduces the per-block partials of launch_segment into
// jac[(j0 + j) * ld + col[s]] = coef[s] * sum_blocks (device jac buffer).
void launch_segment_reduce(Device &dev, int grid, int nslots, int nj, int j0,
                           const int *cols, const double *coefs, double *d_jac,
                           int ld);

// LDS tile passes (kernels_tile.hip).
constexpr int TILE_BLOCK = 256;
constexpr int TILE_TMAX = 14;          // tile bits
constexpr int TILE_PMAX = 24;          // trainable ops per pass
constexpr int TILE_ROWS = TILE_BLOCK / 16; // 16-lane rows holding FP64 partials
constexpr size_t TILE_LDS_MAX = 80 * 1024;

// A bundle: ops[first .. first+nops) on tile-local bits qa (> qb when K == 2);
// its trainable ops use pass slots slot0 .. slot0+nslots.
struct TileBundle {
    int32_t K, qa, qb, first, nops, nslots, slot0, pad;
};

struct TileLaunch {
    int n;
    int T;                // tile bits
    int tb[TILE_TMAX];    // ascending global bit positions of the tile bits
    int nbundles;
    int nslots;           // trainable ops of the pass
    int nj;
    int write_phi;
};

size_t tile_lds_bytes(int dtype, int T, int nj, int nslots);
bool tile_nj_supported(int dtype, int nj);
int launch_tile(Device &dev, const State &phi, void *const *h_ptrs, const void *d_bundles,
                const void *d_ops, const TileLaunch &L);
void launch_tile_reduce(Device &dev, int grid, int nslots, int nj, int j0, const int *cols,
                        const double *coefs, double *d_jac, int ld);

// Register-tile passes (kernels_rtile.hip). A wavefront holds a tile of
// 2^(R+6) amplitudes of phi and nj observable states in VGPRs: R register
// bits (3 for complex128, 4 for complex64) and 6 lane bits, each holding one
// tile bit (a global index bit). A pass program is a list of RtInstr
// followed by the payloads they point to.
constexpr int RT_LANE_BITS = 6;
constexpr int RT_RMAX = 4;
constexpr int RT_TMAX = RT_RMAX + RT_LANE_BITS;
constexpr int RT_PMAX = 32; // trainable ops per pass

enum RtOpCode : int8_t {
    RT_REG1 = 1,  // one wire on register bit a
    RT_REG2 = 2,  // two wires on register bits a > b (local index 2 bit(a) + bit(b))
    RT_SEL = 3,   // register bit a mixes; blocks selected per lane by lane bits s0 (, s1)
    RT_LANE = 4   // diagonal on lane positions s0 (, s1): one factor per lane
};

// Payloads (reals of the state's precision, from the program start):
//   REG1  U: DIAG [u0 u1] complex, REAL [m00 m01 m10 m11], DENSE 2x2 complex,
//            SWAP none; G: DIAG [g0 g1], DENSE 2x2 complex, X / Y / Z none.
//   REG2  U: DIAG 4 complex, REAL 4x4, DENSE 4x4 complex, SWAP none (swap =
//            4 a + b exchanged local indices); G: DIAG 4, DENSE 4x4 complex, Z none.
//   SEL   nb = 2 (s1 < 0) or 4 blocks, block c = bit(s0) + 2 bit(s1), each
//            block as REG1 (U SWAP: one real, nonzero = exchange); G: DIAG or DENSE.
//   LANE  U: nb complex factors; G: nb real factors (same block index).
// Before its operation an instruction applies nremap (0-2) remaps: register
// bit rq exchanged with lane position rp (4 or 5).
struct RtInstr {
    int8_t code, a, b, s0, s1, ukind, gkind, swap;
    int32_t slot; // trainable slot of the pass, -1 otherwise
    int32_t uoff, goff;
    int8_t nremap, rq0, rp0, rq1;
    int8_t rp1, pad8[3];
    int32_t pad;
};
static_assert(sizeof(RtInstr) == 32, "RtInstr layout");

struct RtLaunch {
    int n, T, R;
    int ninstr;
    int nslots;
    int nj;        // observable states per chunk
    int write_phi; // the last chunk writes the updated phi (to phi_out)
    int nc;        // observable chunks run side by side in this launch (>= 1)
    int nbig;      // the first nbig chunks hold nj states, the others nj - 1
    int xgroup;    // block ids taken in groups of xgroup (8: one per XCD) when mapping to chunks
    int tb[RT_TMAX];                   // tile bits, ascending global positions
    int reg_in[RT_RMAX], lane_in[RT_LANE_BITS];   // global bit of each slot at load
    int reg_out[RT_RMAX], lane_out[RT_LANE_BITS]; // and at store
    int nfin;                          // remaps before the store (0-2):
    int fin_rq[2], fin_rp[2];          // register bit fin_rq[f] <-> lane position fin_rp[f]
};

// Runs one register-tile pass over phi and nc chunks of observable states
// side by side: nbig chunks of nj states, then nc - nbig chunks of nj - 1
// (h_ptrs holds their pointers in chunk order). Every chunk
// reads phi; the last one writes the updated phi to phi_out (nullptr: in
// place, allowed only with nc == 1). FP64 partials per block go to dev.red
// as [slot * width + j][block], width = the launch's observable states;
// returns the blocks per chunk (the reduction's grid).
constexpr int RT_OBS_MAX = 24; // observable states per register-tile launch
int launch_rtile(Device &dev, const State &phi, const State *phi_out, void *const *h_ptrs,
                 const void *d_prog, const RtLaunch &L);
// Single-operation steps (kernels_step.hip) for operations the fused passes
// do not hold. One launch applies U to phi (when write_phi) and to nj
// observable states and, for a trainable op, writes the FP64 block sums of
// Im<h_j | G phi> (phi before U) to dev.red as [j][block]; d_mats holds U then
// G restricted to the subspace, D x D FP64 complex matrices (row-major in
// the order of off). Returns the grid size.
constexpr int STEP_MMAX = 4;
struct StepLaunch {
    int m;         // 1..4: U (and G) on a subspace of the op's wires; 0: parity phases
    int nj;        // observable states, 0..SEG_NJMAX
    int write_phi;
    int trainable;
    int dim;       // subspace states D (2, 4, 8, 16)
    int bits[STEP_MMAX];           // ascending global bits of the wires
    uint64_t groups;               // 2^(n - m); parity: 2^n
    uint64_t off[1 << STEP_MMAX];  // global offset of subspace state k from the group base
    uint64_t umask, gmask;         // parity: global bit masks of U's and G's wires
    double ue[2], uo[2];           // parity: U's phase for even / odd parity
};
int launch_step(Device &dev, const State &phi, void *const *h_ptrs, const void *d_mats,
                const StepLaunch &L);

constexpr int RT_NJMAX = 5; // observable states per register-tile launch
// One variant per precision (float / double) and chunk 0..RT_NJMAX, each
// compiled from kernels_rtile.hip as an object of its own.
template <typename T, int NJ>
int rtile_launch_t(Device &dev, const State &phi, void *phi_out, void *const *h_ptrs,
                   const void *d_prog, const RtLaunch &L);

// Batched circuits (kernels_batch.hip). One workgroup runs one circuit at a
// time from |0...0>: forward passes, h = H phi and <H>, reverse passes with
// the FP64 Im<h|G phi> of every trainable op. A pass is a set of T tile bits
// and bundles of ops on <= 2 tile positions. States of n <= tile_bits qubits
// stay in LDS (resident); larger ones live in a per-workgroup slice of a
// global buffer and are streamed through LDS one 2^T-amplitude tile at a time.
constexpr int BATCH_BLOCK = 256;
constexpr int BATCH_PMAX = 64;  // trainable ops per pass
constexpr int BATCH_TMAX = 12;  // tile bits
constexpr int BATCH_NMAX = 20;  // qubits per circuit

enum BatchPassFlags { BP_INIT = 1, BP_NO_STORE = 2 };

struct BatchPass {
    int32_t T;                 // tile bits (== n for a resident circuit)
    int32_t flags;             // BatchPassFlags
    int32_t bundle_first, bundle_count;
    int32_t nslots;            // trainable ops of the pass
    int32_t slot_first;        // their Jacobian columns / coefficients
    uint32_t tmask;            // mask of the tile bits
    int32_t pad;
    int32_t tb[BATCH_TMAX];    // ascending global bit of each tile position
};

struct BatchCircuit {
    int32_t n;
    int32_t resident;
    int32_t fwd_first, fwd_count; // passes
    int32_t bwd_first, bwd_count;
    int32_t term_first, term_count;
    int64_t jac_offset;           // first Jacobian entry of the circuit
    int64_t pad;
};

struct BatchTerm {
    uint64_t x, zy; // Pauli word: flipped bits, signed bits
    double cr, ci;  // coefficient, including (-i)^{#Y}
};

struct BatchArgs {
    const BatchCircuit *circuits;
    const int32_t *order;        // circuit indices, grouped by workgroup
    const int32_t *block_first;  // grid + 1 offsets into order
    const int64_t *block_scratch; // amplitude offset of each workgroup's slice
    const BatchPass *passes;
    const TileBundle *bundles;
    const void *ops;             // packed op groups (OpHead)
    const BatchTerm *terms;
    const int32_t *slot_cols;
    const double *slot_coefs;
    void *scratch;               // Cx<T> amplitudes
    double *expvals;
    double *jac;
    int64_t jac_len;
    int32_t tile_bits;
    int32_t grid;
};

// Parameter patch of a resident batch plan: the U payload of one packed op at
// byte `off` of the op table is rewritten from params[param] (negated when
// adj) by the gate's formula, in the payload layout its ukind implies
// (RX dense, RY real, RZ and PhaseShift diagonal). Circuits whose structure
// repeats between calls (a training loop) reuse the plan and upload only
// their parameters.
enum BatchPatchGate : int32_t { BPG_NONE = 0, BPG_RX = 1, BPG_RY = 2, BPG_RZ = 3, BPG_PHASE = 4 };
struct BatchPatch {
    int64_t off;
    int32_t param;
    int16_t gate; // BatchPatchGate
    int16_t adj;
};
void launch_batch_patch(Device &dev, int dtype, void *ops, const BatchPatch *patches,
                        int64_t count, const double *params);

size_t batch_lds_bytes(int dtype, int tile_bits);
// Workgroups of the batch kernel resident per CU at this tile width.
int batch_resident_blocks(int dtype, int tile_bits);
void launch_batch(Device &dev, int dtype, const BatchArgs &A);

void launch_dense(Device &dev, const State &s, const CompiledOp &op);
void launch_dense_matrix(Device &dev, const State &s,
                         const std::vector<size_t> &wires,
                         const std::vector<cd> &mat);
void launch_parity(Device &dev, const State &s, const std::vector<size_t> &wires,
                   cd even, cd odd);

void launch_fill_zero(Device &dev, const State &s);
// Zeroes len doubles with a kernel (no memset node when a caller captures the stream).
void launch_fill_zero_doubles(Device &dev, double *p, size_t len);
void launch_set_basis(Device &dev, const State &s, size_t index);
void launch_scatter(Device &dev, const State &s, const size_t *d_idx,
                    const void *d_vals, size_t count);
void launch_copy(Device &dev, const State &dst, const State &src);
void launch_axpy(Device &dev, const State &y, double alpha, const State &x);

// sum_i conj(x_j[i]) y[i] for j < nx, FP64; results (re, im) pairs in host out.
void inner_products(Device &dev, const State *const *xs, int nx, const State &y,
                    double *out_re_im);

// Pauli-word sums: terms grouped by X mask.
struct PauliTerm {
    uint64_t x = 0;  // bits flipped (X or Y)
    uint64_t zy = 0; // bits with a sign (Z or Y)
    cd coeff{0, 0};  // includes (-i)^{#Y}
};
void launch_pauli_apply(Device &dev, const State &out, const State &in,
                        const std::vector<PauliTerm> &terms);
double pauli_expval(Device &dev, const State &s,
                    const std::vector<PauliTerm> &terms);
// Device copy of a Pauli-word sum's grouped term tables. Observables keep one
// per (device, qubit count) for their life, so repeated expectation values
// and adjoint sweeps upload nothing.
struct PauliDevice {
    int device = -1;
    hipStream_t stream = nullptr;
    void *blob = nullptr;
    size_t toff = 0; // byte offset of the term table
    int ngroups = 0;
    PauliDevice() = default;
    PauliDevice(const PauliDevice &) = delete;
    PauliDevice &operator=(const PauliDevice &) = delete;
    ~PauliDevice();
};
std::unique_ptr<PauliDevice> pauli_upload(Device &dev, const std::vector<PauliTerm> &terms);
double pauli_expval(Device &dev, const State &s, const PauliDevice &t);
void launch_pauli_apply(Device &dev, const State &out, const State &in, const PauliDevice &t);
// Host image of a Pauli-word sum's tables (the layout PauliDevice holds on
// the device) and an application from tables already at `blob`.
std::vector<unsigned char> pauli_image(const std::vector<PauliTerm> &terms, size_t &toff,
                                       int &ngroups);
void launch_pauli_apply_at(Device &dev, const State &out, const State &in, const void *blob,
                           size_t toff, int ngroups);

// Largest number of Pauli words pauli_expval_noalloc accepts.
constexpr size_t kNoAllocMaxPauliWords = 4096;
// pauli_expval with every device buffer (partials, total, term tables) in a
// fixed file-scope device array instead of allocated memory; raises
// PLK_ERR_INVALID for more than kNoAllocMaxPauliWords terms.
double pauli_expval_noalloc(Device &dev, const State &s, const std::vector<PauliTerm> &terms);

// A CSR matrix as handed over at the boundary: borrowed host arrays, values
// interleaved FP64 complex, indices and indptr both int32 or both int64.
struct CsrView {
    const double *data = nullptr;
    const void *indices = nullptr;
    const void *indptr = nullptr;
    bool idx32 = false;
    size_t nnz = 0;
    size_t rows = 0;
};
// A device copy of a CSR matrix kept for the life of its observable: values
// in the state's precision, 32-bit indices whenever rows and nnz fit.
struct CsrDevice {
    int device = -1;
    hipStream_t stream = nullptr;
    int dtype = 0;
    bool idx32 = true;
    size_t nnz = 0, rows = 0;
    void *blob = nullptr;
    const void *data = nullptr, *indices = nullptr, *indptr = nullptr;
    CsrDevice() = default;
    CsrDevice(const CsrDevice &) = delete;
    CsrDevice &operator=(const CsrDevice &) = delete;
    ~CsrDevice();
};
// O(1) checks of the shape (rows == state length, indptr[0] == 0,
// indptr[rows] == nnz); the kernels check every index and row range and
// report violations, so no O(nnz) host pass is made.
void csr_check_shape(const CsrView &m, size_t len);
std::unique_ptr<CsrDevice> csr_upload(Device &dev, const CsrView &m, int dtype);
// sum_r Re(conj(psi_r) sum_j A_rj psi_cj) (MeasuresKokkos.hpp:132-157).
double csr_expval(Device &dev, const State &s, const CsrDevice &m);
// The same for arrays handed over per call: copied into the device
// workspace (no allocation once it has grown), FP64 values.
double csr_expval_once(Device &dev, const State &s, const CsrView &m);
// out = A in (LinearAlgebraKokkos.hpp:68-148); synchronizes to report
// malformed matrices.
void csr_spmv(Device &dev, const State &out, const State &in, const CsrDevice &m);

// Dense matrices on up to kDenseMaxWires wires (kernels_dense.hip):
// <psi|M|psi> in FP64, and out = M in (out distinct from in).
constexpr size_t kDenseMaxWires = 12;
double dense_expval(Device &dev, const State &s, const std::vector<size_t> &wires,
                    const std::vector<cd> &mat);
void dense_apply_into(Device &dev, const State &out, const State &in,
                      const std::vector<size_t> &wires, const std::vector<cd> &mat);

// Dense 5-qubit blocks on the matrix cores (kernels_block.hip, fuse.cpp).
// A fused block: five ascending state bits and a 32x32 complex matrix over
// the local index s = sum_t bit(bits[t]) << t (row-major).
struct FusedBlock {
    int bits[5];
    std::vector<cd> M;
    int nops = 0; // operations multiplied into M
};
// Greedy fusion of an op list (in order) into blocks of at most five wires:
// a block starts at the first op not yet fused and takes every later op
// whose wires fit in five, skipping (and blocking the wires of) ops that do
// not; ops commute past skipped ops only when their wires are disjoint.
// Blocks with fewer wires are widened with untouched low bits.
std::vector<FusedBlock> fuse_blocks(const std::vector<CompiledOp> &ops, int n);
// The 64 x 64 per-lane operand layout of block5_kernel for M (FP32 for
// complex64 states, FP64 for complex128 states).
int block5_lane_pack(const std::vector<cd> &M, std::vector<float> &out);
int block5_lane_pack(const std::vector<cd> &M, std::vector<double> &out);
// One block pass (n >= 9); d_apack = block5_lane_pack in the state's precision.
void launch_block5(Device &dev, const State &s, const int *bits, const void *d_apack);
// Uploads the lane packs of every block (one copy) and applies them in order.
void apply_blocks(Device &dev, const State &s, const std::vector<FusedBlock> &blocks);

// Probabilities and sampling (kernels_sample.hip). probs: 2^m FP64 values,
// bin bit (m-1-k) taken from global bit gbits[k]. sample: shots rows of n
// bits, wire 0 first.
void probs(Device &dev, const State &s, const std::vector<int> &gbits, double *host_out);
void sample(Device &dev, const State &s, size_t shots, uint64_t seed, size_t *host_out);

} // namespace plk

// Opaque handle types of the C ABI.
struct plk_state {
    plk::State s;
};

namespace plk {
void profiling_enable_impl(bool on);
KernelStats profiling_get_impl(const std::string &kernel);
std::string profiling_names_impl();
} // namespace plk
