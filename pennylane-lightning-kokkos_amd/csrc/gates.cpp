// Host-side gate library: named gates and generators compiled to FP64
// matrices that the kernels consume.
//
// Semantics follow pennylane_lightning_kokkos/src/simulator/GateFunctors.hpp
// (gate functors 302-2170 and 3134-3184; generator functors 2172-3132) and the
// generator scales of StateVectorKokkos.hpp:1275-1580. Matrices use the local
// index r = sum_p bit(wires[p]) << (m-1-p) (GateFunctors.hpp:96-122). The
// inverse flag of a named gate negates its angle or conjugates its phase;
// Rot/CRot map (phi, theta, omega) to (-omega, -theta, -phi)
// (GateFunctors.hpp:805-807, 3151-3153); for an explicit matrix it means the
// conjugate transpose (GateFunctors.hpp:40-53).
#include <algorithm>
#include <cmath>
#include <functional>
#include <unordered_map>
#include <unordered_set>

#include "plk_internal.hpp"

namespace plk {

namespace {

const cd kI{0.0, 1.0};

using Builder = std::function<std::vector<cd>(const std::vector<double> &)>;

struct GateDef {
    size_t wires;  // 0 = any number (MultiRZ)
    size_t params;
    Builder build; // matrix for the given (already sign-adjusted) params
};

std::vector<cd> diag_mat(std::initializer_list<cd> d) {
    const size_t dim = d.size();
    std::vector<cd> M(dim * dim, 0.0);
    size_t i = 0;
    for (auto v : d) {
        M[i * dim + i] = v;
        i++;
    }
    return M;
}

// |1><1| (x) U on the control-1 block of a 4x4 identity.
std::vector<cd> controlled(const std::vector<cd> &u) {
    std::vector<cd> M = diag_mat({1, 1, 1, 1});
    M[10] = u[0];
    M[11] = u[1];
    M[14] = u[2];
    M[15] = u[3];
    return M;
}

std::vector<cd> rx(double t) {
    const double c = std::cos(t / 2), s = std::sin(t / 2);
    return {c, -kI * s, -kI * s, c};
}
std::vector<cd> ry(double t) {
    const double c = std::cos(t / 2), s = std::sin(t / 2);
    return {c, -s, s, c};
}
std::vector<cd> rz(double t) {
    return {std::exp(-kI * (t / 2)), 0.0, 0.0, std::exp(kI * (t / 2))};
}
// Rot(phi, theta, omega) = RZ(omega) RY(theta) RZ(phi).
std::vector<cd> rot(double phi, double theta, double omega) {
    const double c = std::cos(theta / 2), s = std::sin(theta / 2);
    return {std::exp(-kI * ((phi + omega) / 2)) * c,
            -std::exp(kI * ((phi - omega) / 2)) * s,
            std::exp(-kI * ((phi - omega) / 2)) * s,
            std::exp(kI * ((phi + omega) / 2)) * c};
}

// Rotation by theta/2 mixing local indices a and b of a dim x dim block,
// identity phase `e` elsewhere: |a> -> c|a> + s|b>, |b> -> -s|a> + c|b>.
std::vector<cd> givens(size_t dim, size_t a, size_t b, double t, cd e) {
    std::vector<cd> M(dim * dim, 0.0);
    for (size_t i = 0; i < dim; i++) {
        M[i * dim + i] = e;
    }
    const double c = std::cos(t / 2), s = std::sin(t / 2);
    M[a * dim + a] = c;
    M[a * dim + b] = -s;
    M[b * dim + a] = s;
    M[b * dim + b] = c;
    return M;
}

const std::unordered_map<std::string, GateDef> &gate_table() {
    static const std::unordered_map<std::string, GateDef> table = [] {
        std::unordered_map<std::string, GateDef> t;
        auto P = [](const std::vector<double> &p) { return p.at(0); };
        t["PauliX"] = {1, 0, [](auto &) { return std::vector<cd>{0, 1, 1, 0}; }};
        t["PauliY"] = {1, 0, [](auto &) { return std::vector<cd>{0, -kI, kI, 0}; }};
        t["PauliZ"] = {1, 0, [](auto &) { return diag_mat({1, -1}); }};
        t["Hadamard"] = {1, 0, [](auto &) {
                             const double h = M_SQRT1_2;
                             return std::vector<cd>{h, h, h, -h};
                         }};
        // S and T take their sign from the pseudo-parameter +-1 injected below.
        t["S"] = {1, 0, [](auto &p) { return diag_mat({1, p[0] > 0 ? kI : -kI}); }};
        t["T"] = {1, 0, [](auto &p) {
                      return diag_mat({1, std::exp(kI * (p[0] * M_PI / 4))});
                  }};
        t["PhaseShift"] = {1, 1, [P](auto &p) { return diag_mat({1, std::exp(kI * P(p))}); }};
        t["RX"] = {1, 1, [P](auto &p) { return rx(P(p)); }};
        t["RY"] = {1, 1, [P](auto &p) { return ry(P(p)); }};
        t["RZ"] = {1, 1, [P](auto &p) { return rz(P(p)); }};
        t["Rot"] = {1, 3, [](auto &p) { return rot(p.at(0), p.at(1), p.at(2)); }};
        t["CNOT"] = {2, 0, [](auto &) { return controlled({0, 1, 1, 0}); }};
        t["CY"] = {2, 0, [](auto &) { return controlled({0, -kI, kI, 0}); }};
        t["CZ"] = {2, 0, [](auto &) { return diag_mat({1, 1, 1, -1}); }};
        t["SWAP"] = {2, 0, [](auto &) {
                         std::vector<cd> M(16, 0.0);
                         M[0] = M[6] = M[9] = M[15] = 1.0;
                         return M;
                     }};
        t["ControlledPhaseShift"] = {
            2, 1, [P](auto &p) { return diag_mat({1, 1, 1, std::exp(kI * P(p))}); }};
        t["CRX"] = {2, 1, [P](auto &p) { return controlled(rx(P(p))); }};
        t["CRY"] = {2, 1, [P](auto &p) { return controlled(ry(P(p))); }};
        t["CRZ"] = {2, 1, [P](auto &p) { return controlled(rz(P(p))); }};
        t["CRot"] = {2, 3, [](auto &p) { return controlled(rot(p.at(0), p.at(1), p.at(2))); }};
        t["IsingXX"] = {2, 1, [P](auto &p) {
                            const double c = std::cos(P(p) / 2), s = std::sin(P(p) / 2);
                            std::vector<cd> M(16, 0.0);
                            M[0] = M[5] = M[10] = M[15] = c;
                            M[3] = M[6] = M[9] = M[12] = -kI * s;
                            return M;
                        }};
        t["IsingYY"] = {2, 1, [P](auto &p) {
                            const double c = std::cos(P(p) / 2), s = std::sin(P(p) / 2);
                            std::vector<cd> M(16, 0.0);
                            M[0] = M[5] = M[10] = M[15] = c;
                            M[3] = M[12] = kI * s;
                            M[6] = M[9] = -kI * s;
                            return M;
                        }};
        t["IsingXY"] = {2, 1, [P](auto &p) {
                            const double c = std::cos(P(p) / 2), s = std::sin(P(p) / 2);
                            std::vector<cd> M(16, 0.0);
                            M[0] = M[15] = 1.0;
                            M[5] = M[10] = c;
                            M[6] = M[9] = kI * s;
                            return M;
                        }};
        t["IsingZZ"] = {2, 1, [P](auto &p) {
                            const cd a = std::exp(-kI * (P(p) / 2)), b = std::exp(kI * (P(p) / 2));
                            return diag_mat({a, b, b, a});
                        }};
        t["SingleExcitation"] = {2, 1, [P](auto &p) { return givens(4, 1, 2, P(p), 1.0); }};
        t["SingleExcitationMinus"] = {2, 1, [P](auto &p) {
                                          auto M = givens(4, 1, 2, P(p), 1.0);
                                          M[0] = M[15] = std::exp(-kI * (P(p) / 2));
                                          return M;
                                      }};
        t["SingleExcitationPlus"] = {2, 1, [P](auto &p) {
                                         auto M = givens(4, 1, 2, P(p), 1.0);
                                         M[0] = M[15] = std::exp(kI * (P(p) / 2));
                                         return M;
                                     }};
        t["CSWAP"] = {3, 0, [](auto &) {
                          std::vector<cd> M(64, 0.0);
                          for (int i = 0; i < 8; i++) {
                              const int j = (i == 5) ? 6 : (i == 6) ? 5 : i;
                              M[i * 8 + j] = 1.0;
                          }
                          return M;
                      }};
        t["Toffoli"] = {3, 0, [](auto &) {
                            std::vector<cd> M(64, 0.0);
                            for (int i = 0; i < 8; i++) {
                                const int j = (i == 6) ? 7 : (i == 7) ? 6 : i;
                                M[i * 8 + j] = 1.0;
                            }
                            return M;
                        }};
        t["DoubleExcitation"] = {4, 1, [P](auto &p) { return givens(16, 3, 12, P(p), 1.0); }};
        t["DoubleExcitationMinus"] = {4, 1, [P](auto &p) {
                                          return givens(16, 3, 12, P(p), std::exp(-kI * (P(p) / 2)));
                                      }};
        t["DoubleExcitationPlus"] = {4, 1, [P](auto &p) {
                                         return givens(16, 3, 12, P(p), std::exp(kI * (P(p) / 2)));
                                     }};
        t["MultiRZ"] = {0, 1, nullptr};
        t["Identity"] = {0, 0, nullptr};
        return t;
    }();
    return table;
}

struct GenDef {
    size_t wires;
    double scale;
    std::vector<cd> G; // empty for MultiRZ (parity)
};

// Generators as applied in place by the reference's generator functors.
const std::unordered_map<std::string, GenDef> &gen_table() {
    static const std::unordered_map<std::string, GenDef> table = [] {
        std::unordered_map<std::string, GenDef> t;
        auto m4 = [](std::initializer_list<std::tuple<int, int, cd>> e) {
            std::vector<cd> M(16, 0.0);
            for (auto &[r, c, v] : e) {
                M[r * 4 + c] = v;
            }
            return M;
        };
        t["RX"] = {1, -0.5, {0, 1, 1, 0}};
        t["RY"] = {1, -0.5, {0, -kI, kI, 0}};
        t["RZ"] = {1, -0.5, {1, 0, 0, -1}};
        t["PhaseShift"] = {1, 1.0, {0, 0, 0, 1}};
        t["ControlledPhaseShift"] = {2, 1.0, m4({{3, 3, 1.0}})};
        t["CRX"] = {2, -0.5, m4({{2, 3, 1.0}, {3, 2, 1.0}})};
        t["CRY"] = {2, -0.5, m4({{2, 3, -kI}, {3, 2, kI}})};
        t["CRZ"] = {2, -0.5, m4({{2, 2, 1.0}, {3, 3, -1.0}})};
        t["IsingXX"] = {2, -0.5, m4({{0, 3, 1.0}, {1, 2, 1.0}, {2, 1, 1.0}, {3, 0, 1.0}})};
        t["IsingYY"] = {2, -0.5, m4({{0, 3, -1.0}, {1, 2, 1.0}, {2, 1, 1.0}, {3, 0, -1.0}})};
        t["IsingZZ"] = {2, -0.5, m4({{0, 0, 1.0}, {1, 1, -1.0}, {2, 2, -1.0}, {3, 3, 1.0}})};
        t["IsingXY"] = {2, 0.5, m4({{1, 2, 1.0}, {2, 1, 1.0}})};
        t["SingleExcitation"] = {2, -0.5, m4({{1, 2, -kI}, {2, 1, kI}})};
        t["SingleExcitationMinus"] = {
            2, -0.5, m4({{0, 0, 1.0}, {1, 2, -kI}, {2, 1, kI}, {3, 3, 1.0}})};
        t["SingleExcitationPlus"] = {
            2, -0.5, m4({{0, 0, -1.0}, {1, 2, -kI}, {2, 1, kI}, {3, 3, -1.0}})};
        auto de = [](bool identity_rest, bool plus) {
            std::vector<cd> M(256, 0.0);
            if (identity_rest) {
                for (int i = 0; i < 16; i++) {
                    M[i * 16 + i] = 1.0;
                }
            }
            M[3 * 16 + 3] = 0.0;
            M[12 * 16 + 12] = 0.0;
            M[3 * 16 + 12] = plus ? kI : -kI;
            M[12 * 16 + 3] = plus ? -kI : kI;
            return M;
        };
        t["DoubleExcitation"] = {4, -0.5, de(false, false)};
        t["DoubleExcitationMinus"] = {4, -0.5, de(true, false)};
        t["DoubleExcitationPlus"] = {4, 0.5, de(true, true)};
        t["MultiRZ"] = {0, -0.5, {}};
        return t;
    }();
    return table;
}

void check_wires(const std::vector<size_t> &wires, int n) {
    std::unordered_set<size_t> seen;
    for (auto w : wires) {
        if (w >= static_cast<size_t>(n)) {
            raise(PLK_ERR_INVALID, "wire index " + std::to_string(w) +
                                       " out of range for " + std::to_string(n) +
                                       " qubits");
        }
        if (!seen.insert(w).second) {
            raise(PLK_ERR_INVALID, "wires of an operation must be distinct");
        }
    }
}

} // namespace

bool is_named_gate(const std::string &name) {
    return gate_table().count(name) > 0;
}

size_t named_gate_wires(const std::string &name) {
    return gate_table().at(name).wires;
}

size_t named_gate_params(const std::string &name) {
    return gate_table().at(name).params;
}

bool has_generator(const std::string &name) { return gen_table().count(name) > 0; }

CompiledOp compile_op(const std::string &name, const std::vector<size_t> &wires,
                      bool inverse, const std::vector<double> &params,
                      const std::vector<cd> &matrix, int num_qubits) {
    CompiledOp op;
    op.name = name;
    op.wires = wires;
    check_wires(wires, num_qubits);
    if (name == "Identity") {
        op.kind = OpKind::Identity;
        return op;
    }
    auto it = gate_table().find(name);
    if (it != gate_table().end()) {
        const GateDef &g = it->second;
        if (g.wires != 0 && wires.size() != g.wires) {
            raise(PLK_ERR_INVALID, "gate " + name + " acts on " +
                                       std::to_string(g.wires) + " wires, got " +
                                       std::to_string(wires.size()));
        }
        if (params.size() < g.params) {
            raise(PLK_ERR_INVALID, "gate " + name + " needs " +
                                       std::to_string(g.params) + " parameters");
        }
        if (name == "MultiRZ") {
            if (wires.empty()) {
                raise(PLK_ERR_INVALID, "MultiRZ needs at least one wire");
            }
            const double t = inverse ? -params[0] : params[0];
            op.kind = OpKind::Parity;
            op.even = std::exp(-kI * (t / 2));
            op.odd = std::exp(kI * (t / 2));
            return op;
        }
        std::vector<double> p;
        if (name == "S" || name == "T") {
            p = {inverse ? -1.0 : 1.0};
        } else if (name == "Rot" || name == "CRot") {
            p = inverse ? std::vector<double>{-params[2], -params[1], -params[0]}
                        : std::vector<double>{params[0], params[1], params[2]};
        } else {
            for (size_t i = 0; i < g.params; i++) {
                p.push_back(inverse ? -params[i] : params[i]);
            }
        }
        op.kind = OpKind::Dense;
        op.mat = g.build(p);
        return op;
    }
    // Unnamed operation: explicit matrix (StateVectorKokkos.hpp:595-599).
    const size_t dim = size_t{1} << wires.size();
    if (wires.empty() || matrix.size() != dim * dim) {
        raise(PLK_ERR_UNSUPPORTED,
              "operation " + name + " has no kernel and no " +
                  std::to_string(dim) + "x" + std::to_string(dim) + " matrix");
    }
    op.kind = OpKind::Dense;
    op.mat.resize(dim * dim);
    for (size_t r = 0; r < dim; r++) {
        for (size_t c = 0; c < dim; c++) {
            op.mat[r * dim + c] = inverse ? std::conj(matrix[c * dim + r])
                                          : matrix[r * dim + c];
        }
    }
    return op;
}

CompiledGen compile_generator(const std::string &name,
                              const std::vector<size_t> &wires, int num_qubits) {
    auto it = gen_table().find(name);
    if (it == gen_table().end()) {
        raise(PLK_ERR_INVALID, "Generator does not exist for " + name);
    }
    check_wires(wires, num_qubits);
    const GenDef &g = it->second;
    CompiledGen out;
    out.wires = wires;
    out.scale = g.scale;
    if (g.wires == 0) {
        out.kind = OpKind::Parity;
        return out;
    }
    if (wires.size() != g.wires) {
        raise(PLK_ERR_INVALID, "generator of " + name + " acts on " +
                                   std::to_string(g.wires) + " wires");
    }
    out.kind = OpKind::Dense;
    out.mat = g.G;
    return out;
}

} // namespace plk
