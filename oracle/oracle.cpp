// CPU oracle for the adjoint-Jacobian state-vector path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, in plain C++17 with
// std::complex<double> and OpenMP, the algorithms of PennyLane-Lightning-Kokkos
// 0.33.0-dev for the hot path (gates, generators, observables, expectation
// values and the adjoint Jacobian sweep). It is linked only by the tests, by
// the smoke check and by the benchmark's CPU-baseline leg, always as the
// checker or as the timed CPU baseline, never as the product.
//
// Conventions restated from the reference (paths relative to
// pennylane_lightning_kokkos/src/):
//   * wire w lives at bit (n-1-w) of the amplitude index
//     (simulator/GateFunctors.hpp:32, rev_wire = num_qubits - wires[0] - 1);
//   * a k-wire operation's local index is r = sum_p bit(wires[p]) << (k-1-p),
//     i.e. wires[0] is the most significant local bit
//     (simulator/GateFunctors.hpp:96-122 and 175-189);
//   * state storage is interleaved (re, im) complex<double>.
//
// Parity pins (checked by tests/test_oracle_*.py):
//   * named gates: golden vectors transcribed from
//     src/tests/Test_StateVectorKokkos_Param.cpp (IsingXX/XY/YY/ZZ, MultiRZ,
//     RX, RY, RZ, PhaseShift, ControlledPhaseShift, Rot, CRot, Single- and
//     DoubleExcitation+-) and _NonParam.cpp (Hadamard, PauliX/Y/Z, S, T,
//     CNOT, SWAP, CZ, Toffoli, CSWAP); every gate (CY, CRX, CRY, CRZ
//     included) against a textbook matrix written in the test from its
//     definition (controlled gates as |0><0| I + |1><1| U), and the inverse
//     flag as U^dagger;
//   * generators: numpy matrices of G as the reference's generator kernels
//     apply it, with their scales, written in the test; U(theta) ==
//     expm(i s theta G) from those two independent tables; central finite
//     differences of the gate (src/tests/Test_StateVectorKokkos_Generator.cpp);
//   * expectation values: goldens of src/tests/Test_StateVectorKokkos_Expval.cpp
//     and tests/test_hamiltonian*.py, plus dense numpy brute force;
//   * adjoint Jacobian: goldens of src/tests/Test_AdjointDiffKokkos.cpp, central
//     finite differences and exact parameter shift on random circuits;
//   * CSR SpMV: goldens of src/tests/Test_LinearAlgebra.cpp;
//   * probabilities: goldens of src/tests/Test_StateVectorKokkos_Measure.cpp
//     and tests/test_probs.py, plus numpy marginalisation of |psi|^2;
//   * sampling (index selection for given uniforms): the bracketing property
//     cdf[k] < u <= cdf[k+1] and basis states.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using cd = std::complex<double>;
using Mat = std::vector<cd>; // row-major 2^m x 2^m

struct OracleError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

static thread_local std::string g_last_error;

[[noreturn]] static void fail(const std::string &msg) { throw OracleError(msg); }

// ---------------------------------------------------------------------------
// Index helpers (util/UtilKokkos.hpp:97-106 fillTrailingOnes/fillLeadingOnes;
// the reference inserts zero bits with parity masks, restated here as an
// ascending-order zero-bit insertion).
// ---------------------------------------------------------------------------
static inline size_t insert_zero(size_t k, int b) {
    const size_t low = k & ((size_t{1} << b) - 1);
    return ((k >> b) << (b + 1)) | low;
}

struct Layout {
    int n = 0;
    int m = 0;
    std::vector<int> sorted_bits;  // ascending bit positions of the wires
    std::vector<size_t> offsets;   // offset of local index r
    size_t groups = 0;
};

static Layout make_layout(int n, const std::vector<size_t> &wires) {
    Layout L;
    L.n = n;
    L.m = static_cast<int>(wires.size());
    std::unordered_set<size_t> seen;
    for (auto w : wires) {
        if (w >= static_cast<size_t>(n)) {
            fail("wire index out of range");
        }
        if (!seen.insert(w).second) {
            fail("wires must be distinct");
        }
        L.sorted_bits.push_back(n - 1 - static_cast<int>(w));
    }
    std::sort(L.sorted_bits.begin(), L.sorted_bits.end());
    const size_t dim = size_t{1} << L.m;
    L.offsets.resize(dim);
    for (size_t r = 0; r < dim; r++) {
        size_t off = 0;
        for (int p = 0; p < L.m; p++) {
            if ((r >> (L.m - 1 - p)) & 1U) {
                off |= size_t{1} << (n - 1 - static_cast<int>(wires[p]));
            }
        }
        L.offsets[r] = off;
    }
    L.groups = size_t{1} << (n - L.m);
    return L;
}

static inline size_t group_base(const Layout &L, size_t k) {
    for (int b : L.sorted_bits) {
        k = insert_zero(k, b);
    }
    return k;
}

// Apply a dense 2^m x 2^m matrix (row-major) to every group.
// Restates singleQubitOpFunctor / twoQubitOpFunctor / multiQubitOpFunctor
// (simulator/GateFunctors.hpp:15-300); the inverse flag means conjugate
// transpose (GateFunctors.hpp:40-53, 116-160, 238-260).
static void apply_dense(cd *sv, int n, const std::vector<size_t> &wires,
                        const Mat &mat, bool inverse) {
    const Layout L = make_layout(n, wires);
    const size_t dim = size_t{1} << L.m;
    if (mat.size() != dim * dim) {
        fail("matrix size does not match the number of wires");
    }
    Mat M = mat;
    if (inverse) {
        for (size_t i = 0; i < dim; i++) {
            for (size_t j = 0; j < dim; j++) {
                M[i * dim + j] = std::conj(mat[j * dim + i]);
            }
        }
    }
    const long long G = static_cast<long long>(L.groups);
#pragma omp parallel
    {
        std::vector<cd> v(dim);
#pragma omp for schedule(static)
        for (long long k = 0; k < G; k++) {
            const size_t base = group_base(L, static_cast<size_t>(k));
            for (size_t r = 0; r < dim; r++) {
                v[r] = sv[base + L.offsets[r]];
            }
            for (size_t r = 0; r < dim; r++) {
                cd acc = 0.0;
                for (size_t c = 0; c < dim; c++) {
                    acc += M[r * dim + c] * v[c];
                }
                sv[base + L.offsets[r]] = acc;
            }
        }
    }
}

// Apply a diagonal (one complex factor per local index) to every group.
static void apply_diag(cd *sv, int n, const std::vector<size_t> &wires,
                       const std::vector<cd> &d) {
    const Layout L = make_layout(n, wires);
    const size_t dim = size_t{1} << L.m;
    const long long G = static_cast<long long>(L.groups);
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < G; k++) {
        const size_t base = group_base(L, static_cast<size_t>(k));
        for (size_t r = 0; r < dim; r++) {
            sv[base + L.offsets[r]] *= d[r];
        }
    }
}

// ---------------------------------------------------------------------------
// Named gates (simulator/GateFunctors.hpp:302-2170, 3134-3184; dispatch table
// simulator/StateVectorKokkos.hpp:141-344). Each gate is restated as its
// dense matrix; the inverse flag follows the reference's rule for named gates
// (negated angle / conjugated phase, Rot and CRot map (phi,theta,omega) to
// (-omega,-theta,-phi): GateFunctors.hpp:805-807, 3151-3153).
// ---------------------------------------------------------------------------
static const cd I1{0.0, 1.0};

static Mat eye(size_t dim) {
    Mat M(dim * dim, 0.0);
    for (size_t i = 0; i < dim; i++) {
        M[i * dim + i] = 1.0;
    }
    return M;
}

static Mat rot_matrix(double phi, double theta, double omega) {
    // Rot(phi, theta, omega) = RZ(omega) RY(theta) RZ(phi)
    // (GateFunctors.hpp:3134-3184 rotFunctor; GatesHost.hpp:433-458 getRot).
    const double c = std::cos(theta / 2);
    const double s = std::sin(theta / 2);
    const double p = phi + omega;
    const double mm = phi - omega;
    return {std::exp(-I1 * (p / 2)) * c, -std::exp(I1 * (mm / 2)) * s,
            std::exp(-I1 * (mm / 2)) * s, std::exp(I1 * (p / 2)) * c};
}

static void embed_controlled(Mat &M4, const Mat &U2) {
    // |1><1| (x) U2 on the control-1 block (local indices 2, 3).
    M4[2 * 4 + 2] = U2[0];
    M4[2 * 4 + 3] = U2[1];
    M4[3 * 4 + 2] = U2[2];
    M4[3 * 4 + 3] = U2[3];
}

static size_t gate_num_params(const std::string &name) {
    static const std::unordered_map<std::string, size_t> np{
        {"Identity", 0}, {"PauliX", 0}, {"PauliY", 0}, {"PauliZ", 0},
        {"Hadamard", 0}, {"S", 0}, {"T", 0}, {"CNOT", 0}, {"SWAP", 0},
        {"CSWAP", 0}, {"Toffoli", 0}, {"CY", 0}, {"CZ", 0},
        {"PhaseShift", 1}, {"ControlledPhaseShift", 1}, {"RX", 1}, {"RY", 1},
        {"RZ", 1}, {"Rot", 3}, {"CRX", 1}, {"CRY", 1}, {"CRZ", 1},
        {"CRot", 3}, {"IsingXX", 1}, {"IsingXY", 1}, {"IsingYY", 1},
        {"IsingZZ", 1}, {"MultiRZ", 1}, {"SingleExcitation", 1},
        {"SingleExcitationMinus", 1}, {"SingleExcitationPlus", 1},
        {"DoubleExcitation", 1}, {"DoubleExcitationMinus", 1},
        {"DoubleExcitationPlus", 1}};
    auto it = np.find(name);
    if (it == np.end()) {
        return SIZE_MAX;
    }
    return it->second;
}

static bool is_named_gate(const std::string &name) {
    return gate_num_params(name) != SIZE_MAX;
}

static size_t gate_num_wires(const std::string &name) {
    static const std::unordered_map<std::string, size_t> nw{
        {"PauliX", 1}, {"PauliY", 1}, {"PauliZ", 1}, {"Hadamard", 1},
        {"S", 1}, {"T", 1}, {"PhaseShift", 1}, {"RX", 1}, {"RY", 1},
        {"RZ", 1}, {"Rot", 1}, {"CNOT", 2}, {"SWAP", 2}, {"CY", 2},
        {"CZ", 2}, {"ControlledPhaseShift", 2}, {"CRX", 2}, {"CRY", 2},
        {"CRZ", 2}, {"CRot", 2}, {"IsingXX", 2}, {"IsingXY", 2},
        {"IsingYY", 2}, {"IsingZZ", 2}, {"SingleExcitation", 2},
        {"SingleExcitationMinus", 2}, {"SingleExcitationPlus", 2},
        {"CSWAP", 3}, {"Toffoli", 3}, {"DoubleExcitation", 4},
        {"DoubleExcitationMinus", 4}, {"DoubleExcitationPlus", 4}};
    auto it = nw.find(name);
    return it == nw.end() ? 0 : it->second;
}

// Dense matrix of a named gate with the reference's inverse rule applied.
static Mat named_gate_matrix(const std::string &name,
                             const std::vector<double> &params, bool inv) {
    const double sgn = inv ? -1.0 : 1.0;
    auto ang = [&](size_t i) {
        if (i >= params.size()) {
            fail("missing parameter for gate " + name);
        }
        return params[i];
    };
    if (name == "PauliX") {
        return {0, 1, 1, 0};
    }
    if (name == "PauliY") {
        return {0, -I1, I1, 0};
    }
    if (name == "PauliZ") {
        return {1, 0, 0, -1};
    }
    if (name == "Hadamard") {
        const double h = M_SQRT1_2;
        return {h, h, h, -h};
    }
    if (name == "S") {
        return {1, 0, 0, inv ? -I1 : I1};
    }
    if (name == "T") {
        const cd t = std::exp(I1 * (M_PI / 4));
        return {1, 0, 0, inv ? std::conj(t) : t};
    }
    if (name == "PhaseShift") {
        return {1, 0, 0, std::exp(I1 * (sgn * ang(0)))};
    }
    if (name == "RX") {
        const double c = std::cos(sgn * ang(0) / 2);
        const double s = std::sin(sgn * ang(0) / 2);
        return {c, -I1 * s, -I1 * s, c};
    }
    if (name == "RY") {
        const double c = std::cos(sgn * ang(0) / 2);
        const double s = std::sin(sgn * ang(0) / 2);
        return {c, -s, s, c};
    }
    if (name == "RZ") {
        const double t = sgn * ang(0) / 2;
        return {std::exp(-I1 * t), 0, 0, std::exp(I1 * t)};
    }
    if (name == "Rot") {
        if (inv) {
            return rot_matrix(-ang(2), -ang(1), -ang(0));
        }
        return rot_matrix(ang(0), ang(1), ang(2));
    }
    if (name == "CNOT") {
        Mat M = eye(4);
        embed_controlled(M, {0, 1, 1, 0});
        return M;
    }
    if (name == "CY") {
        Mat M = eye(4);
        embed_controlled(M, {0, -I1, I1, 0});
        return M;
    }
    if (name == "CZ") {
        Mat M = eye(4);
        M[15] = -1.0;
        return M;
    }
    if (name == "SWAP") {
        Mat M(16, 0.0);
        M[0] = M[1 * 4 + 2] = M[2 * 4 + 1] = M[15] = 1.0;
        return M;
    }
    if (name == "ControlledPhaseShift") {
        Mat M = eye(4);
        M[15] = std::exp(I1 * (sgn * ang(0)));
        return M;
    }
    if (name == "CRX" || name == "CRY" || name == "CRZ") {
        Mat M = eye(4);
        embed_controlled(M, named_gate_matrix(name.substr(1), params, inv));
        return M;
    }
    if (name == "CRot") {
        Mat M = eye(4);
        embed_controlled(M, named_gate_matrix("Rot", params, inv));
        return M;
    }
    if (name == "IsingXX" || name == "IsingYY" || name == "IsingXY" ||
        name == "IsingZZ" || name == "SingleExcitation" ||
        name == "SingleExcitationMinus" || name == "SingleExcitationPlus") {
        const double th = sgn * ang(0);
        const double c = std::cos(th / 2);
        const double s = std::sin(th / 2);
        Mat M(16, 0.0);
        auto at = [&](int r, int col) -> cd & { return M[r * 4 + col]; };
        if (name == "IsingXX") { // GateFunctors.hpp:895-959
            at(0, 0) = at(1, 1) = at(2, 2) = at(3, 3) = c;
            at(0, 3) = at(1, 2) = at(2, 1) = at(3, 0) = -I1 * s;
        } else if (name == "IsingYY") { // GateFunctors.hpp:1025-1088
            at(0, 0) = at(1, 1) = at(2, 2) = at(3, 3) = c;
            at(0, 3) = at(3, 0) = I1 * s;
            at(1, 2) = at(2, 1) = -I1 * s;
        } else if (name == "IsingXY") { // GateFunctors.hpp:961-1023
            at(0, 0) = at(3, 3) = 1.0;
            at(1, 1) = at(2, 2) = c;
            at(1, 2) = at(2, 1) = I1 * s;
        } else if (name == "IsingZZ") { // GateFunctors.hpp:1090-1151
            const cd e0 = std::exp(-I1 * (th / 2));
            const cd e1 = std::exp(I1 * (th / 2));
            at(0, 0) = at(3, 3) = e0;
            at(1, 1) = at(2, 2) = e1;
        } else { // SingleExcitation(+/-), GateFunctors.hpp:1153-1340
            cd e = 1.0;
            if (name == "SingleExcitationMinus") {
                e = std::exp(-I1 * (th / 2));
            } else if (name == "SingleExcitationPlus") {
                e = std::exp(I1 * (th / 2));
            }
            at(0, 0) = at(3, 3) = e;
            at(1, 1) = at(2, 2) = c;
            at(1, 2) = -s;
            at(2, 1) = s;
        }
        return M;
    }
    if (name == "CSWAP") { // GateFunctors.hpp:1995-2061: swap |101>,|110>
        Mat M = eye(8);
        M[5 * 8 + 5] = M[6 * 8 + 6] = 0.0;
        M[5 * 8 + 6] = M[6 * 8 + 5] = 1.0;
        return M;
    }
    if (name == "Toffoli") { // GateFunctors.hpp:2063-2130: swap |110>,|111>
        Mat M = eye(8);
        M[6 * 8 + 6] = M[7 * 8 + 7] = 0.0;
        M[6 * 8 + 7] = M[7 * 8 + 6] = 1.0;
        return M;
    }
    if (name == "DoubleExcitation" || name == "DoubleExcitationMinus" ||
        name == "DoubleExcitationPlus") { // GateFunctors.hpp:1342-1766
        const double th = sgn * ang(0);
        const double c = std::cos(th / 2);
        const double s = std::sin(th / 2);
        cd e = 1.0;
        if (name == "DoubleExcitationMinus") {
            e = std::exp(-I1 * (th / 2));
        } else if (name == "DoubleExcitationPlus") {
            e = std::exp(I1 * (th / 2));
        }
        Mat M(256, 0.0);
        for (int r = 0; r < 16; r++) {
            M[r * 16 + r] = e;
        }
        M[3 * 16 + 3] = c;
        M[3 * 16 + 12] = -s;
        M[12 * 16 + 3] = s;
        M[12 * 16 + 12] = c;
        return M;
    }
    fail("unknown gate " + name);
}

// MultiRZ is a diagonal on an arbitrary number of wires
// (GateFunctors.hpp:2132-2170): phase e^{-i theta/2} for even parity of the
// local index, e^{+i theta/2} for odd parity.
static std::vector<cd> multirz_diag(size_t m, double theta) {
    std::vector<cd> d(size_t{1} << m);
    const cd e0 = std::exp(-I1 * (theta / 2));
    const cd e1 = std::exp(I1 * (theta / 2));
    for (size_t r = 0; r < d.size(); r++) {
        d[r] = (__builtin_popcountll(r) % 2 == 0) ? e0 : e1;
    }
    return d;
}

// StateVectorKokkos::applyOperation (StateVectorKokkos.hpp:585-600):
// "Identity" is a no-op, named gates go through the dispatch table, anything
// else is applied as a dense matrix (inverse = conjugate transpose).
static void apply_operation(cd *sv, int n, const std::string &name,
                            const std::vector<size_t> &wires, bool inverse,
                            const std::vector<double> &params,
                            const Mat &matrix) {
    if (name == "Identity") {
        return;
    }
    if (name == "MultiRZ") {
        if (params.empty()) {
            fail("missing parameter for gate MultiRZ");
        }
        const double th = inverse ? -params[0] : params[0];
        apply_diag(sv, n, wires, multirz_diag(wires.size(), th));
        return;
    }
    if (is_named_gate(name)) {
        if (wires.size() != gate_num_wires(name)) {
            fail("wrong number of wires for gate " + name);
        }
        apply_dense(sv, n, wires, named_gate_matrix(name, params, inverse),
                    false);
        return;
    }
    if (matrix.empty()) {
        fail("operation " + name + " has no kernel and no matrix");
    }
    apply_dense(sv, n, wires, matrix, inverse);
}

// ---------------------------------------------------------------------------
// Generators (simulator/GateFunctors.hpp:2172-3132; scales
// simulator/StateVectorKokkos.hpp:1275-1580; RX/RY/RZ reuse PauliX/Y/Z with
// scale -1/2, StateVectorKokkos.hpp:1452-1488). G is applied in place and the
// scale s with U(theta) = exp(i s theta G) is returned. The generator's
// adjoint flag is ignored by the reference functors.
// ---------------------------------------------------------------------------
struct Generator {
    Mat G;          // dense generator on the op's wires (empty for MultiRZ)
    double scale;
};

static Generator generator_of(const std::string &name, size_t nwires) {
    auto sq = [](std::initializer_list<std::pair<int, std::pair<int, cd>>> e,
                 int dim) {
        Mat M(static_cast<size_t>(dim * dim), 0.0);
        for (auto &kv : e) {
            M[kv.first * dim + kv.second.first] = kv.second.second;
        }
        return M;
    };
    if (name == "RX") {
        return {{0, 1, 1, 0}, -0.5};
    }
    if (name == "RY") {
        return {{0, -I1, I1, 0}, -0.5};
    }
    if (name == "RZ") {
        return {{1, 0, 0, -1}, -0.5};
    }
    if (name == "PhaseShift") { // zero |0> component
        return {{0, 0, 0, 1}, 1.0};
    }
    if (name == "ControlledPhaseShift") { // |11><11|
        return {sq({{3, {3, 1.0}}}, 4), 1.0};
    }
    if (name == "CRX") {
        return {sq({{2, {3, 1.0}}, {3, {2, 1.0}}}, 4), -0.5};
    }
    if (name == "CRY") {
        return {sq({{2, {3, -I1}}, {3, {2, I1}}}, 4), -0.5};
    }
    if (name == "CRZ") {
        return {sq({{2, {2, 1.0}}, {3, {3, -1.0}}}, 4), -0.5};
    }
    if (name == "IsingXX") {
        return {sq({{0, {3, 1.0}}, {3, {0, 1.0}}, {1, {2, 1.0}}, {2, {1, 1.0}}},
                   4),
                -0.5};
    }
    if (name == "IsingYY") {
        return {sq({{0, {3, -1.0}}, {3, {0, -1.0}}, {1, {2, 1.0}}, {2, {1, 1.0}}},
                   4),
                -0.5};
    }
    if (name == "IsingZZ") {
        return {sq({{0, {0, 1.0}}, {1, {1, -1.0}}, {2, {2, -1.0}}, {3, {3, 1.0}}},
                   4),
                -0.5};
    }
    if (name == "IsingXY") {
        return {sq({{1, {2, 1.0}}, {2, {1, 1.0}}}, 4), 0.5};
    }
    if (name == "SingleExcitation") {
        return {sq({{1, {2, -I1}}, {2, {1, I1}}}, 4), -0.5};
    }
    if (name == "SingleExcitationMinus") {
        return {sq({{0, {0, 1.0}}, {1, {2, -I1}}, {2, {1, I1}}, {3, {3, 1.0}}},
                   4),
                -0.5};
    }
    if (name == "SingleExcitationPlus") {
        return {sq({{0, {0, -1.0}}, {1, {2, -I1}}, {2, {1, I1}}, {3, {3, -1.0}}},
                   4),
                -0.5};
    }
    if (name == "DoubleExcitation" || name == "DoubleExcitationMinus" ||
        name == "DoubleExcitationPlus") {
        Mat M(256, 0.0);
        const bool plus = name == "DoubleExcitationPlus";
        if (name != "DoubleExcitation") {
            for (int r = 0; r < 16; r++) {
                M[r * 16 + r] = 1.0;
            }
        }
        M[3 * 16 + 3] = 0.0;
        M[12 * 16 + 12] = 0.0;
        M[3 * 16 + 12] = plus ? I1 : -I1;
        M[12 * 16 + 3] = plus ? -I1 : I1;
        return {M, plus ? 0.5 : -0.5};
    }
    if (name == "MultiRZ") {
        (void)nwires;
        return {{}, -0.5};
    }
    fail("Generator does not exist for " + name);
}

static double apply_generator(cd *sv, int n, const std::string &name,
                              const std::vector<size_t> &wires) {
    Generator g = generator_of(name, wires.size());
    if (name == "MultiRZ") {
        std::vector<cd> d(size_t{1} << wires.size());
        for (size_t r = 0; r < d.size(); r++) {
            d[r] = (__builtin_popcountll(r) % 2 == 0) ? 1.0 : -1.0;
        }
        apply_diag(sv, n, wires, d);
        return g.scale;
    }
    apply_dense(sv, n, wires, g.G, false);
    return g.scale;
}

// ---------------------------------------------------------------------------
// Linear algebra (util/LinearAlgebraKokkos.hpp).
// ---------------------------------------------------------------------------
static double real_inner(const cd *x, const cd *y, size_t N) { // 155-192
    double acc = 0.0;
    const long long NN = static_cast<long long>(N);
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (long long i = 0; i < NN; i++) {
        acc += std::real(std::conj(x[i]) * y[i]);
    }
    return acc;
}

static double imag_inner(const cd *x, const cd *y, size_t N) { // 199-236
    double acc = 0.0;
    const long long NN = static_cast<long long>(N);
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (long long i = 0; i < NN; i++) {
        acc += std::imag(std::conj(x[i]) * y[i]);
    }
    return acc;
}

static void axpy(cd alpha, const cd *x, cd *y, size_t N) { // 30-61
    const long long NN = static_cast<long long>(N);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < NN; i++) {
        y[i] += alpha * x[i];
    }
}

static void copy_state(cd *dst, const cd *src, size_t N) {
    const long long NN = static_cast<long long>(N);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < NN; i++) {
        dst[i] = src[i];
    }
}

// y = A x with A in CSR (LinearAlgebraKokkos.hpp:68-148).
static void spmv(const cd *x, cd *y, const std::vector<cd> &data,
                 const std::vector<int64_t> &indices,
                 const std::vector<int64_t> &indptr) {
    const long long rows = static_cast<long long>(indptr.size()) - 1;
#pragma omp parallel for schedule(static)
    for (long long r = 0; r < rows; r++) {
        cd acc = 0.0;
        for (int64_t j = indptr[r]; j < indptr[r + 1]; j++) {
            acc += data[j] * x[indices[j]];
        }
        y[r] = acc;
    }
}

// ---------------------------------------------------------------------------
// Observables (simulator/ObservablesKokkos.hpp:19-517).
// ---------------------------------------------------------------------------
struct Obs {
    enum Kind { Named, Hermitian, Tensor, Hamiltonian, Sparse } kind;
    std::string name;
    std::vector<size_t> wires;
    Mat matrix;
    std::vector<std::shared_ptr<Obs>> terms;
    std::vector<double> coeffs;
    std::vector<cd> data;
    std::vector<int64_t> indices, indptr;
};

static void obs_apply(cd *sv, int n, const Obs &ob);

static void obs_apply(cd *sv, int n, const Obs &ob) {
    const size_t N = size_t{1} << n;
    switch (ob.kind) {
    case Obs::Named: // NamedObsKokkos::applyInPlace, 116-118
        apply_operation(sv, n, ob.name, ob.wires, false, {}, {});
        return;
    case Obs::Hermitian: // HermitianObsKokkos::applyInPlace (multi-qubit op)
        apply_dense(sv, n, ob.wires, ob.matrix, false);
        return;
    case Obs::Tensor: // TensorProdObsKokkos::applyInPlace, 272-276
        for (const auto &t : ob.terms) {
            obs_apply(sv, n, *t);
        }
        return;
    case Obs::Hamiltonian: { // HamiltonianKokkos::applyInPlace, 360-373
        std::vector<cd> buffer(N, 0.0);
        std::vector<cd> tmp(N);
        for (size_t t = 0; t < ob.terms.size(); t++) {
            copy_state(tmp.data(), sv, N);
            obs_apply(tmp.data(), n, *ob.terms[t]);
            axpy(cd{ob.coeffs[t], 0.0}, tmp.data(), buffer.data(), N);
        }
        copy_state(sv, buffer.data(), N);
        return;
    }
    case Obs::Sparse: { // SparseHamiltonianKokkos::applyInPlace, 484-494
        if (ob.wires.size() != static_cast<size_t>(n)) {
            fail("SparseH wire count does not match state-vector size");
        }
        std::vector<cd> out(N);
        spmv(sv, out.data(), ob.data, ob.indices, ob.indptr);
        copy_state(sv, out.data(), N);
        return;
    }
    }
}

// MeasuresKokkos::expval(ob) (MeasuresKokkos.hpp:354-360).
static double obs_expval(const cd *sv, int n, const Obs &ob) {
    const size_t N = size_t{1} << n;
    std::vector<cd> tmp(sv, sv + N);
    obs_apply(tmp.data(), n, ob);
    return real_inner(sv, tmp.data(), N);
}

// ---------------------------------------------------------------------------
// Operations list (algorithms/AdjointDiffKokkos.hpp:17-173).
// ---------------------------------------------------------------------------
struct Ops {
    std::vector<std::string> names;
    std::vector<std::vector<double>> params;
    std::vector<std::vector<size_t>> wires;
    std::vector<bool> inverses;
    std::vector<Mat> matrices;
    size_t num_par_ops() const {
        size_t c = 0;
        for (const auto &p : params) {
            c += p.empty() ? 0 : 1;
        }
        return c;
    }
};

// AdjointJacobianKokkos<T>::adjointJacobian (AdjointDiffKokkos.hpp:404-478),
// restated step by step, including the per-op copy of mu and the serial
// loops over observables.
static void adjoint_jacobian(const cd *ref, int n,
                             const std::vector<std::shared_ptr<Obs>> &obs,
                             const Ops &ops, const std::vector<size_t> &tp,
                             bool apply_operations, double *jac) {
    if (tp.empty()) {
        fail("No trainable parameters provided.");
    }
    const size_t N = size_t{1} << n;
    const size_t n_obs = obs.size();
    const size_t tp_size = tp.size();
    const size_t num_param_ops = ops.num_par_ops();
    std::fill(jac, jac + n_obs * tp_size, 0.0);

    std::vector<cd> lambda(ref, ref + N);
    if (apply_operations) { // applyOperations, 213-226
        for (size_t i = 0; i < ops.names.size(); i++) {
            apply_operation(lambda.data(), n, ops.names[i], ops.wires[i],
                            ops.inverses[i], ops.params[i], ops.matrices[i]);
        }
    }
    std::vector<std::vector<cd>> H(n_obs, std::vector<cd>(N));
    for (size_t j = 0; j < n_obs; j++) { // applyObservables, 262-293
        copy_state(H[j].data(), lambda.data(), N);
        obs_apply(H[j].data(), n, *obs[j]);
    }
    std::vector<cd> mu(N);

    size_t trainable_param_number = tp_size - 1;
    size_t current_param_idx = num_param_ops - 1;
    auto tp_it = tp.rbegin();
    for (long op = static_cast<long>(ops.names.size()) - 1; op >= 0; op--) {
        const size_t o = static_cast<size_t>(op);
        if (ops.params[o].size() > 1) {
            fail("The operation is not supported using the adjoint "
                 "differentiation method");
        }
        if (ops.names[o] == "StatePrep" || ops.names[o] == "BasisState") {
            continue;
        }
        if (tp_it == tp.rend()) {
            break;
        }
        copy_state(mu.data(), lambda.data(), N);
        apply_operation(lambda.data(), n, ops.names[o], ops.wires[o],
                        !ops.inverses[o], ops.params[o], ops.matrices[o]);
        if (!ops.params[o].empty()) {
            if (current_param_idx == *tp_it) {
                const double scale =
                    apply_generator(mu.data(), n, ops.names[o], ops.wires[o]) *
                    (ops.inverses[o] ? -1.0 : 1.0);
                for (size_t j = 0; j < n_obs; j++) {
                    jac[j * tp_size + trainable_param_number] =
                        -2.0 * scale * imag_inner(H[j].data(), mu.data(), N);
                }
                trainable_param_number--;
                ++tp_it;
            }
            current_param_idx--;
        }
        for (size_t j = 0; j < n_obs; j++) {
            apply_operation(H[j].data(), n, ops.names[o], ops.wires[o],
                            !ops.inverses[o], ops.params[o], ops.matrices[o]);
        }
    }
}

} // namespace orc

// ---------------------------------------------------------------------------
// C interface used through ctypes by tests/ and the benchmark's CPU leg.
// Return code 0 = success, 1 = error (message in orc_last_error()).
// ---------------------------------------------------------------------------
using namespace orc;

#define ORC_TRY(body)                                                          \
    try {                                                                      \
        body;                                                                  \
        return 0;                                                              \
    } catch (const std::exception &e) {                                        \
        g_last_error = e.what();                                               \
        return 1;                                                              \
    }

static std::vector<size_t> to_wires(const int64_t *w, int nw) {
    std::vector<size_t> out(static_cast<size_t>(nw));
    for (int i = 0; i < nw; i++) {
        if (w[i] < 0) {
            fail("negative wire index");
        }
        out[static_cast<size_t>(i)] = static_cast<size_t>(w[i]);
    }
    return out;
}

static Mat to_mat(const double *m, size_t len) {
    Mat M(len);
    for (size_t i = 0; i < len; i++) {
        M[i] = cd{m[2 * i], m[2 * i + 1]};
    }
    return M;
}

// ---------------------------------------------------------------------------
// Probabilities and sampling (simulator/MeasuresKokkos.hpp:389-569,
// simulator/MeasuresFunctors.hpp:15-210, util/UtilKokkos.hpp:353-581).
// ---------------------------------------------------------------------------

// generateBitsPatterns (UtilKokkos.hpp:567-583): the amplitude offsets of the
// 2^k assignments of `qubits`, the first qubit most significant.
std::vector<size_t> bit_patterns(const std::vector<size_t> &qubits, int n) {
    std::vector<size_t> idx{0};
    for (size_t it0 = 0; it0 < qubits.size(); it0++) {
        const size_t q = qubits[qubits.size() - 1 - it0];
        const size_t value = size_t{1} << (n - 1 - static_cast<int>(q));
        const size_t cur = idx.size();
        for (size_t j = 0; j < cur; j++) {
            idx.push_back(idx[j] + value);
        }
    }
    return idx;
}

// MeasuresKokkos::probs(wires) (MeasuresKokkos.hpp:413-520): sub-probabilities
// over the sorted wires, then the transposition to the given wire order
// (getTransposedIndexFunctor, MeasuresFunctors.hpp:144-175).
std::vector<double> probs_of(const cd *sv, int n, const std::vector<size_t> &wires) {
    std::vector<size_t> order(wires.size());
    for (size_t k = 0; k < order.size(); k++) {
        order[k] = k;
    }
    std::sort(order.begin(), order.end(),
              [&](size_t a, size_t b) { return wires[a] < wires[b]; });
    std::vector<size_t> sorted(wires.size());
    for (size_t k = 0; k < order.size(); k++) {
        sorted[k] = wires[order[k]];
    }
    std::vector<size_t> rest;
    for (int q = 0; q < n; q++) {
        if (std::find(sorted.begin(), sorted.end(), static_cast<size_t>(q)) == sorted.end()) {
            rest.push_back(static_cast<size_t>(q));
        }
    }
    const auto idx = bit_patterns(sorted, n);
    const auto off = bit_patterns(rest, n);
    std::vector<double> sub(idx.size(), 0.0);
    for (size_t i = 0; i < idx.size(); i++) {
        for (size_t j = 0; j < off.size(); j++) {
            sub[i] += std::norm(sv[idx[i] + off[j]]);
        }
    }
    const size_t m = wires.size();
    std::vector<double> out(idx.size());
    for (size_t i = 0; i < idx.size(); i++) {
        size_t t = 0;
        for (size_t j = 0; j < m; j++) {
            const size_t bit = (i >> (m - 1 - j)) & 1;
            t += bit << (m - 1 - order[j]);
        }
        out[i] = sub[t];
    }
    return out;
}

// generate_samples (MeasuresKokkos.hpp:533-569) with the uniforms given:
// exclusive CDF (getCDFFunctor) and the binary search of Sampler
// (MeasuresFunctors.hpp:74-110). Returns basis indices.
void sample_indices(const cd *sv, int n, const double *u, size_t shots, int64_t *out) {
    const size_t N = size_t{1} << n;
    std::vector<double> cdf(N);
    double run = 0.0;
    for (size_t k = 0; k < N; k++) {
        cdf[k] = run;
        run += std::norm(sv[k]);
    }
    for (size_t s = 0; s < shots; s++) {
        size_t index;
        if (u[s] <= cdf[1]) {
            index = 0;
        } else {
            size_t low = 1, high = N;
            while (high - low > 1) {
                const size_t mid = high - ((high - low) >> 1);
                const double c = mid == N ? 1.0 : cdf[mid];
                if (c < u[s]) {
                    low = mid;
                } else {
                    high = mid;
                }
            }
            index = high - 1;
        }
        out[s] = static_cast<int64_t>(index);
    }
}

extern "C" {

const char *orc_last_error(void) { return g_last_error.c_str(); }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_apply(double *sv, int n, const char *name, const int64_t *wires,
              int nw, int inverse, const double *params, int np,
              const double *mat, int64_t mat_len) {
    ORC_TRY({
        std::vector<double> p(params, params + np);
        Mat M = mat ? to_mat(mat, static_cast<size_t>(mat_len)) : Mat{};
        apply_operation(reinterpret_cast<cd *>(sv), n, name,
                        to_wires(wires, nw), inverse != 0, p, M);
    })
}

int orc_apply_matrix(double *sv, int n, const double *mat,
                     const int64_t *wires, int nw, int inverse) {
    ORC_TRY({
        const size_t dim = size_t{1} << nw;
        apply_dense(reinterpret_cast<cd *>(sv), n, to_wires(wires, nw),
                    to_mat(mat, dim * dim), inverse != 0);
    })
}

int orc_apply_generator(double *sv, int n, const char *name,
                        const int64_t *wires, int nw, double *scale) {
    ORC_TRY({
        *scale = apply_generator(reinterpret_cast<cd *>(sv), n, name,
                                 to_wires(wires, nw));
    })
}

// Dense matrix of a named gate (row-major, interleaved complex, 4^m doubles*2).
int orc_gate_matrix(const char *name, int nw, int inverse,
                    const double *params, int np, double *out) {
    ORC_TRY({
        std::vector<double> p(params, params + np);
        Mat M;
        if (std::string(name) == "MultiRZ") {
            auto d = multirz_diag(static_cast<size_t>(nw),
                                  inverse ? -p.at(0) : p.at(0));
            M.assign(d.size() * d.size(), 0.0);
            for (size_t i = 0; i < d.size(); i++) {
                M[i * d.size() + i] = d[i];
            }
        } else {
            M = named_gate_matrix(name, p, inverse != 0);
        }
        std::memcpy(out, M.data(), M.size() * sizeof(cd));
    })
}

// Dense generator of a parametric gate and its scale.
int orc_generator_matrix(const char *name, int nw, double *out,
                         double *scale) {
    ORC_TRY({
        Generator g = generator_of(name, static_cast<size_t>(nw));
        if (std::string(name) == "MultiRZ") {
            const size_t dim = size_t{1} << nw;
            g.G.assign(dim * dim, 0.0);
            for (size_t r = 0; r < dim; r++) {
                g.G[r * dim + r] = (__builtin_popcountll(r) % 2 == 0) ? 1.0 : -1.0;
            }
        }
        std::memcpy(out, g.G.data(), g.G.size() * sizeof(cd));
        *scale = g.scale;
    })
}

int orc_gate_info(const char *name, int *num_params, int *num_wires) {
    ORC_TRY({
        const size_t np = gate_num_params(name);
        if (np == SIZE_MAX) {
            fail(std::string("unknown gate ") + name);
        }
        *num_params = static_cast<int>(np);
        *num_wires = static_cast<int>(gate_num_wires(name));
    })
}

void *orc_obs_named(const char *name, const int64_t *wires, int nw) {
    try {
        auto *o = new Obs{};
        o->kind = Obs::Named;
        o->name = name;
        o->wires = to_wires(wires, nw);
        return o;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void *orc_obs_hermitian(const double *mat, const int64_t *wires, int nw) {
    try {
        auto *o = new Obs{};
        o->kind = Obs::Hermitian;
        o->wires = to_wires(wires, nw);
        const size_t dim = size_t{1} << nw;
        o->matrix = to_mat(mat, dim * dim);
        return o;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

// The factors are copied; the caller keeps ownership of its handles.
void *orc_obs_tensor(void *const *factors, int nf) {
    try {
        auto *o = new Obs{};
        o->kind = Obs::Tensor;
        std::unordered_set<size_t> seen;
        for (int i = 0; i < nf; i++) {
            auto f = std::make_shared<Obs>(*static_cast<Obs *>(factors[i]));
            std::vector<size_t> fw = f->wires;
            if (f->kind == Obs::Tensor || f->kind == Obs::Hamiltonian) {
                for (auto &t : f->terms) {
                    fw.insert(fw.end(), t->wires.begin(), t->wires.end());
                }
            }
            for (auto w : fw) {
                if (!seen.insert(w).second) {
                    fail("All wires in observables must be disjoint.");
                }
            }
            o->terms.push_back(f);
        }
        return o;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void *orc_obs_hamiltonian(const double *coeffs, void *const *terms, int nt) {
    try {
        auto *o = new Obs{};
        o->kind = Obs::Hamiltonian;
        for (int i = 0; i < nt; i++) {
            o->coeffs.push_back(coeffs[i]);
            o->terms.push_back(
                std::make_shared<Obs>(*static_cast<Obs *>(terms[i])));
        }
        return o;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void *orc_obs_sparse(const double *data, const int64_t *indices,
                     const int64_t *indptr, int64_t nnz, int64_t nrows,
                     const int64_t *wires, int nw) {
    try {
        auto *o = new Obs{};
        o->kind = Obs::Sparse;
        o->data = to_mat(data, static_cast<size_t>(nnz));
        o->indices.assign(indices, indices + nnz);
        o->indptr.assign(indptr, indptr + nrows + 1);
        o->wires = to_wires(wires, nw);
        return o;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void orc_obs_free(void *o) { delete static_cast<Obs *>(o); }

int orc_obs_apply(double *sv, int n, const void *o) {
    ORC_TRY({ obs_apply(reinterpret_cast<cd *>(sv), n, *static_cast<const Obs *>(o)); })
}

int orc_expval_obs(const double *sv, int n, const void *o, double *out) {
    ORC_TRY({
        *out = obs_expval(reinterpret_cast<const cd *>(sv), n,
                          *static_cast<const Obs *>(o));
    })
}

// CSR expectation value: sum_r Re(conj(psi_r) sum_j A_rj psi_{c_j})
// (ExpValFunctors.hpp:252-280).
static double csr_expval(const cd *psi, const cd *d, const int64_t *indices,
                         const int64_t *indptr, int64_t nrows) {
    double acc = 0.0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (long long r = 0; r < nrows; r++) {
        for (int64_t j = indptr[r]; j < indptr[r + 1]; j++) {
            acc += std::real(std::conj(psi[r]) * d[j] * psi[indices[j]]);
        }
    }
    return acc;
}

int orc_expval_csr(const double *sv, int n, const double *data,
                   const int64_t *indices, const int64_t *indptr,
                   int64_t nrows, double *out) {
    ORC_TRY({
        (void)n;
        *out = csr_expval(reinterpret_cast<const cd *>(sv),
                          reinterpret_cast<const cd *>(data), indices, indptr,
                          nrows);
    })
}

int orc_spmv(const double *x, double *y, const double *data,
             const int64_t *indices, const int64_t *indptr, int64_t nrows,
             int64_t nnz) {
    ORC_TRY({
        std::vector<cd> dv = to_mat(data, static_cast<size_t>(nnz));
        std::vector<int64_t> iv(indices, indices + nnz);
        std::vector<int64_t> pv(indptr, indptr + nrows + 1);
        spmv(reinterpret_cast<const cd *>(x), reinterpret_cast<cd *>(y), dv,
             iv, pv);
    })
}

int orc_inner(const double *x, const double *y, int64_t N, double *re,
              double *im) {
    ORC_TRY({
        *re = real_inner(reinterpret_cast<const cd *>(x),
                         reinterpret_cast<const cd *>(y), static_cast<size_t>(N));
        *im = imag_inner(reinterpret_cast<const cd *>(x),
                         reinterpret_cast<const cd *>(y), static_cast<size_t>(N));
    })
}

// Operation list. params/wires/matrices are flattened with offsets of length
// n_ops + 1; matrices are interleaved complex (mat_off counts complex entries).
void *orc_ops_create(int n_ops, const char *const *names,
                     const double *params, const int64_t *param_off,
                     const int64_t *wires, const int64_t *wire_off,
                     const int *inverses, const double *mats,
                     const int64_t *mat_off) {
    try {
        auto *ops = new Ops{};
        for (int i = 0; i < n_ops; i++) {
            ops->names.emplace_back(names[i]);
            ops->params.emplace_back(params + param_off[i],
                                     params + param_off[i + 1]);
            ops->wires.push_back(to_wires(wires + wire_off[i],
                                          static_cast<int>(wire_off[i + 1] -
                                                           wire_off[i])));
            ops->inverses.push_back(inverses[i] != 0);
            const int64_t ml = mat_off[i + 1] - mat_off[i];
            ops->matrices.push_back(
                ml > 0 ? to_mat(mats + 2 * mat_off[i], static_cast<size_t>(ml))
                       : Mat{});
        }
        return ops;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void orc_ops_free(void *ops) { delete static_cast<Ops *>(ops); }

int orc_adjoint_jacobian(const double *sv, int n, void *const *obs, int n_obs,
                         const void *ops, const int64_t *tp, int n_tp,
                         int apply_ops, double *jac) {
    ORC_TRY({
        std::vector<std::shared_ptr<Obs>> ob;
        for (int j = 0; j < n_obs; j++) {
            ob.push_back(std::make_shared<Obs>(*static_cast<Obs *>(obs[j])));
        }
        std::vector<size_t> t(static_cast<size_t>(n_tp));
        for (int i = 0; i < n_tp; i++) {
            t[static_cast<size_t>(i)] = static_cast<size_t>(tp[i]);
        }
        adjoint_jacobian(reinterpret_cast<const cd *>(sv), n, ob,
                         *static_cast<const Ops *>(ops), t, apply_ops != 0, jac);
    })
}

int orc_probs(const double *sv, int n, const int64_t *wires, int nw, double *out) {
    ORC_TRY({
        std::vector<size_t> w;
        for (int k = 0; k < nw; k++) {
            w.push_back(static_cast<size_t>(wires[k]));
        }
        if (w.empty()) {
            for (int q = 0; q < n; q++) {
                w.push_back(static_cast<size_t>(q));
            }
        }
        const auto p = probs_of(reinterpret_cast<const cd *>(sv), n, w);
        std::copy(p.begin(), p.end(), out);
    })
}

int orc_sample_indices(const double *sv, int n, const double *uniforms, int64_t shots,
                       int64_t *out) {
    ORC_TRY({
        sample_indices(reinterpret_cast<const cd *>(sv), n, uniforms,
                       static_cast<size_t>(shots), out);
    })
}

} // extern "C"
