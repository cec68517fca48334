// C ABI of the engine: states, gates, observables, measurements and the
// adjoint Jacobian. See include/plk_amd.h for the contract of each entry
// point and the reference member it replaces.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <set>
#include <sstream>
#include <thread>

#include "plk_internal.hpp"

namespace plk {

namespace {

thread_local std::string g_last_error;

const cd kI{0.0, 1.0};

Device &dev_of(const State &s) { return device(s.device); }

// ---------------------------------------------------------------------------
// Observables (ObservablesKokkos.hpp:19-517).
// ---------------------------------------------------------------------------
struct ObsNode {
    enum Kind { Named, Hermitian, Tensor, Hamiltonian, Sparse } kind = Named;
    std::string name;
    std::vector<size_t> wires;
    std::vector<cd> matrix;
    std::vector<std::shared_ptr<const ObsNode>> terms;
    std::vector<double> coeffs;
    std::vector<double> data; // CSR values, interleaved complex
    std::vector<int64_t> indices, indptr;       // int64 matrices
    std::vector<int32_t> indices32, indptr32;   // int32 matrices (complex64 serialisation)
    bool idx32 = false;
    size_t rows = 0;
    // Device copies of the CSR matrix per (device, state dtype), made on first
    // use and kept while the observable lives (entry points hold the engine
    // lock, so the cache needs no lock of its own).
    mutable std::map<std::pair<int, int>, std::unique_ptr<CsrDevice>> csr_dev;
    // Device Pauli tables per (device, qubit count) when the observable is a
    // sum of Pauli words (made on first use, kept while the observable lives).
    mutable std::map<std::pair<int, int>, std::unique_ptr<PauliDevice>> pauli_dev;
    const PauliDevice *pauli_on(Device &dev, int n) const;

    CsrView csr_view() const {
        CsrView v;
        v.data = data.data();
        v.idx32 = idx32;
        v.indices = idx32 ? static_cast<const void *>(indices32.data()) : indices.data();
        v.indptr = idx32 ? static_cast<const void *>(indptr32.data()) : indptr.data();
        v.nnz = data.size() / 2;
        v.rows = rows;
        return v;
    }
    const CsrDevice &csr_on(Device &dev, const State &s) const {
        auto &slot = csr_dev[{dev.id, s.dtype}];
        if (!slot) {
            const CsrView v = csr_view();
            csr_check_shape(v, s.len);
            slot = csr_upload(dev, v, s.dtype);
        }
        return *slot;
    }
};

std::vector<size_t> obs_wires(const ObsNode &o) {
    std::set<size_t> w;
    if (o.kind == ObsNode::Tensor || o.kind == ObsNode::Hamiltonian) {
        for (auto &t : o.terms) {
            auto tw = obs_wires(*t);
            w.insert(tw.begin(), tw.end());
        }
    } else {
        w.insert(o.wires.begin(), o.wires.end());
    }
    return {w.begin(), w.end()};
}

constexpr size_t kMaxPauliTerms = 4096;

uint64_t wire_bit(int n, size_t w) {
    if (w >= static_cast<size_t>(n)) {
        raise(PLK_ERR_INVALID, "observable wire " + std::to_string(w) + " out of range");
    }
    return uint64_t{1} << (n - 1 - static_cast<int>(w));
}

// Expands an observable into Pauli words when it is a sum of products of
// Identity / PauliX / PauliY / PauliZ / Hadamard with at most cap words.
// Returns false otherwise.
bool to_pauli(const ObsNode &o, int n, std::vector<PauliTerm> &out,
              size_t cap = kMaxPauliTerms) {
    switch (o.kind) {
    case ObsNode::Named: {
        if (o.name == "Identity") {
            out = {PauliTerm{0, 0, 1.0}};
            return true;
        }
        if (o.wires.size() != 1) {
            return false;
        }
        const uint64_t b = wire_bit(n, o.wires[0]);
        if (o.name == "PauliX") {
            out = {PauliTerm{b, 0, 1.0}};
        } else if (o.name == "PauliY") {
            out = {PauliTerm{b, b, -kI}};
        } else if (o.name == "PauliZ") {
            out = {PauliTerm{0, b, 1.0}};
        } else if (o.name == "Hadamard") {
            out = {PauliTerm{b, 0, M_SQRT1_2}, PauliTerm{0, b, M_SQRT1_2}};
        } else {
            return false;
        }
        return true;
    }
    case ObsNode::Tensor: {
        std::vector<PauliTerm> acc{PauliTerm{0, 0, 1.0}};
        for (auto &f : o.terms) {
            std::vector<PauliTerm> ft;
            if (!to_pauli(*f, n, ft, cap)) {
                return false;
            }
            std::vector<PauliTerm> next;
            if (acc.size() * ft.size() > cap) {
                return false;
            }
            for (auto &a : acc) {
                for (auto &b : ft) {
                    // Factors act on disjoint wires, so words multiply bitwise.
                    next.push_back({a.x | b.x, a.zy | b.zy, a.coeff * b.coeff});
                }
            }
            acc.swap(next);
        }
        out = acc;
        return true;
    }
    case ObsNode::Hamiltonian: {
        std::vector<PauliTerm> acc;
        for (size_t t = 0; t < o.terms.size(); t++) {
            std::vector<PauliTerm> tt;
            if (!to_pauli(*o.terms[t], n, tt, cap)) {
                return false;
            }
            for (auto &p : tt) {
                p.coeff *= o.coeffs[t];
                acc.push_back(p);
            }
            if (acc.size() > cap) {
                return false;
            }
        }
        out = acc;
        return true;
    }
    default:
        return false;
    }
}

// ---------------------------------------------------------------------------
// Gate application.
// ---------------------------------------------------------------------------
struct PlanOp {
    CompiledOp U;
    bool trainable = false;
    CompiledGen G;
    int col = -1;
    double coef = 0.0;
    int32_t src = -1;       // index of the op in its OpsData
    int32_t param = -1;     // batch plans: global parameter index of a patched op
    int16_t patch_gate = 0; // batch plans: BatchPatchGate of a patched op
    bool adj = false;       // U was compiled with the inverse flag set
};

// Parity ops/generators on <= 2 wires become dense diagonals so they fuse.
std::vector<cd> parity_diag(size_t m, cd even, cd odd) {
    const size_t dim = size_t{1} << m;
    std::vector<cd> M(dim * dim, 0.0);
    for (size_t r = 0; r < dim; r++) {
        M[r * dim + r] = (__builtin_popcountll(r) & 1) ? odd : even;
    }
    return M;
}

void densify(CompiledOp &op) {
    if (op.kind == OpKind::Parity && op.wires.size() <= 2) {
        op.mat = parity_diag(op.wires.size(), op.even, op.odd);
        op.kind = OpKind::Dense;
    }
}

void densify(CompiledGen &g) {
    if (g.kind == OpKind::Parity && g.wires.size() <= 2) {
        g.mat = parity_diag(g.wires.size(), 1.0, -1.0);
        g.kind = OpKind::Dense;
    }
}

bool fusable(const PlanOp &p) {
    if (p.U.kind != OpKind::Dense || p.U.wires.size() > 2) {
        return false;
    }
    if (p.trainable && (p.G.kind != OpKind::Dense || p.G.wires.size() > 2)) {
        return false;
    }
    return true;
}

struct Segment {
    std::vector<size_t> wires; // global wires of the segment
    std::vector<PlanOp> ops;
    int nslots = 0;
};

// A tile pass: bundles of ops on at most two tile bits each, all inside the
// pass's T tile bits.
struct Bundle {
    int K = 1;
    int qa = 0, qb = 0; // tile-local positions, qa > qb when K == 2
    std::vector<PlanOp> ops;
    int nslots = 0;
};

struct Pass {
    std::vector<int> tb; // ascending global bit positions
    std::vector<Bundle> bundles;
    std::vector<PlanOp> ops; // register-tile passes: the ops in program order
    bool rt = false;         // a register-tile pass (ops) rather than LDS bundles
    int nslots = 0;
    std::vector<int> cols;     // Jacobian column of each trainable op, in order
    std::vector<double> coefs; // and its coefficient
};

struct Step {
    bool fused = false; // register segment
    bool tiled = false; // LDS tile pass
    Segment seg;
    Pass pass;
    PlanOp single;
};

constexpr size_t kSegMaxOps = 64;

// Greedy fusion of consecutive operations into register segments.
std::vector<Step> plan_steps(std::vector<PlanOp> &&ops) {
    std::vector<Step> steps;
    Segment cur;
    auto flush = [&]() {
        if (!cur.ops.empty()) {
            Step s;
            s.fused = true;
            s.seg = std::move(cur);
            steps.push_back(std::move(s));
        }
        cur = Segment{};
    };
    for (auto &p : ops) {
        if (p.U.kind == OpKind::Identity && !p.trainable) {
            continue;
        }
        densify(p.U);
        if (p.trainable) {
            densify(p.G);
        }
        if (!fusable(p)) {
            flush();
            Step s;
            s.fused = false;
            s.single = std::move(p);
            steps.push_back(std::move(s));
            continue;
        }
        std::set<size_t> w(cur.wires.begin(), cur.wires.end());
        w.insert(p.U.wires.begin(), p.U.wires.end());
        if (w.size() > static_cast<size_t>(SEG_KMAX) || cur.ops.size() >= kSegMaxOps ||
            (p.trainable && cur.nslots >= SEG_PMAX)) {
            flush();
            w = std::set<size_t>(p.U.wires.begin(), p.U.wires.end());
        }
        cur.wires.assign(w.begin(), w.end());
        if (p.trainable) {
            cur.nslots++;
        }
        cur.ops.push_back(std::move(p));
    }
    flush();
    return steps;
}

void prepare(PlanOp &p) {
    densify(p.U);
    if (p.trainable) {
        densify(p.G);
    }
}

constexpr size_t kPassMaxOps = 512;
constexpr size_t kBundleLookahead = 64;
constexpr size_t kLookahead = 512;

// Splits a pass's ops (in order) into bundles of <= 2 tile bits.
void make_bundles(Pass &ps, std::vector<PlanOp> &&ops, int n) {
    int pos[64];
    for (size_t q = 0; q < ps.tb.size(); q++) {
        pos[ps.tb[q]] = static_cast<int>(q);
    }
    std::vector<PlanOp> all = std::move(ops);
    std::vector<uint32_t> obits(all.size());
    std::vector<uint64_t> wmask(all.size());
    std::vector<int> rem(all.size());
    for (size_t i = 0; i < all.size(); i++) {
        uint32_t ob = 0;
        uint64_t wm = 0;
        for (auto w : all[i].U.wires) {
            ob |= 1u << pos[n - 1 - static_cast<int>(w)];
            wm |= uint64_t{1} << w;
        }
        obits[i] = ob;
        wmask[i] = wm;
        rem[i] = static_cast<int>(i);
    }
    // Each bundle starts at the first remaining op and takes later ops that
    // fit its two tile bits, provided no skipped op before them shares a
    // wire (ops on disjoint wires commute), within a window of kBundleLookahead.
    std::vector<int> rest;
    while (!rem.empty()) {
        Bundle cur;
        uint32_t cb = 0;
        uint64_t blocked = 0;
        rest.clear();
        size_t k = 0;
        for (; k < rem.size() && k < kBundleLookahead; k++) {
            const int i = rem[k];
            PlanOp &p = all[i];
            const bool ok = !(blocked & wmask[i]) && __builtin_popcount(cb | obits[i]) <= 2 &&
                            cur.ops.size() < kSegMaxOps && (!p.trainable || cur.nslots < SEG_PMAX);
            if (ok) {
                cb |= obits[i];
                if (p.trainable) {
                    cur.nslots++;
                    ps.cols.push_back(p.col);
                    ps.coefs.push_back(p.coef);
                    ps.nslots++;
                }
                cur.ops.push_back(std::move(p));
            } else {
                blocked |= wmask[i];
                rest.push_back(i);
            }
        }
        rest.insert(rest.end(), rem.begin() + static_cast<long>(k), rem.end());
        const int lo = __builtin_ctz(cb);
        if (__builtin_popcount(cb) == 1) {
            cur.K = 1;
            cur.qa = lo;
        } else {
            cur.K = 2;
            cur.qa = 31 - __builtin_clz(cb);
            cur.qb = lo;
        }
        ps.bundles.push_back(std::move(cur));
        rem.swap(rest);
    }
}

// True when the dim x dim matrix M has a nonzero entry between local
// indices that differ in local bit lb.
bool mixes_bit(const std::vector<cd> &M, int dim, int lb) {
    for (int r = 0; r < dim; r++) {
        for (int c = 0; c < dim; c++) {
            if ((((r ^ c) >> lb) & 1) && M[static_cast<size_t>(r) * dim + c] != cd(0.0, 0.0)) {
                return true;
            }
        }
    }
    return false;
}

// Global state bits an op mixes (its U or, when trainable, its generator has
// an entry between indices that differ in the bit).
uint64_t mixing_mask(const PlanOp &p, int n) {
    const int m = static_cast<int>(p.U.wires.size());
    uint64_t mask = 0;
    for (int k = 0; k < m; k++) {
        const int lb = m - 1 - k;
        if (mixes_bit(p.U.mat, 1 << m, lb) || (p.trainable && mixes_bit(p.G.mat, 1 << m, lb))) {
            mask |= uint64_t{1} << (n - 1 - static_cast<int>(p.U.wires[k]));
        }
    }
    return mask;
}

// Groups operations into LDS tile passes of T bits that always contain the
// lowest L bits. Operations on disjoint wires commute, so an operation may be
// taken ahead of skipped ones it shares no wire with; an unfusable operation
// is a barrier and runs on its own. With bundled == false (register-tile
// passes) at most maxmix distinct bits may be mixed per pass.
std::vector<Step> plan_tiles(std::vector<PlanOp> &&ops_in, int n, int T, int L,
                             int pmax = TILE_PMAX, bool bundled = true, int maxmix = 64) {
    std::vector<PlanOp> all;
    for (auto &p : ops_in) {
        if (p.U.kind == OpKind::Identity && !p.trainable) {
            continue;
        }
        prepare(p);
        all.push_back(std::move(p));
    }
    // Wire and state-bit masks (wires and bits < 64).
    std::vector<uint64_t> wmask(all.size()), bmask(all.size());
    std::vector<char> fus(all.size());
    std::vector<int> rem(all.size());
    for (size_t i = 0; i < all.size(); i++) {
        for (auto w : all[i].U.wires) {
            wmask[i] |= uint64_t{1} << w;
            bmask[i] |= uint64_t{1} << (n - 1 - static_cast<int>(w));
        }
        fus[i] = fusable(all[i]) ? 1 : 0;
        rem[i] = static_cast<int>(i);
    }
    std::vector<uint64_t> mmask(all.size(), 0);
    if (!bundled) {
        for (size_t i = 0; i < all.size(); i++) {
            if (fus[i]) {
                mmask[i] = mixing_mask(all[i], n);
            }
        }
    }
    std::vector<Step> steps;
    std::vector<int> taken, rest;
    while (!rem.empty()) {
        if (!fus[rem[0]]) {
            Step s;
            s.single = std::move(all[rem[0]]);
            steps.push_back(std::move(s));
            rem.erase(rem.begin());
            continue;
        }
        uint64_t bits = L > 0 ? (uint64_t{1} << L) - 1 : 0;
        uint64_t blocked = 0, mixed = 0;
        taken.clear();
        rest.clear();
        int nslots = 0;
        size_t k = 0;
        for (; k < rem.size(); k++) {
            const int i = rem[k];
            if (!fus[i] || k >= kLookahead) {
                break;
            }
            const bool tr = all[i].trainable;
            // Register-tile passes need only the bits an op mixes in the
            // tile: its other bits may lie outside, where they are fixed per
            // tile and select the op's block.
            const uint64_t nb = bits | (bundled ? bmask[i] : mmask[i]);
            const bool ok = taken.size() < kPassMaxOps && (!tr || nslots < pmax) &&
                            !(blocked & wmask[i]) && __builtin_popcountll(nb) <= T &&
                            __builtin_popcountll(mixed | mmask[i]) <= maxmix;
            if (ok) {
                bits = nb;
                mixed |= mmask[i];
                nslots += tr ? 1 : 0;
                taken.push_back(i);
            } else {
                blocked |= wmask[i];
                rest.push_back(i);
            }
        }
        rest.insert(rest.end(), rem.begin() + static_cast<long>(k), rem.end());
        // Pad the tile with the lowest unused bits.
        for (int b = 0; __builtin_popcountll(bits) < T && b < n; b++) {
            bits |= uint64_t{1} << b;
        }
        Step s;
        s.tiled = true;
        for (int b = 0; b < 64; b++) {
            if (bits >> b & 1) {
                s.pass.tb.push_back(b);
            }
        }
        std::vector<PlanOp> ops;
        ops.reserve(taken.size());
        for (int i : taken) {
            ops.push_back(std::move(all[i]));
        }
        if (bundled) {
            make_bundles(s.pass, std::move(ops), n);
        } else {
            for (auto &p : ops) {
                if (p.trainable) {
                    s.pass.cols.push_back(p.col);
                    s.pass.coefs.push_back(p.coef);
                    s.pass.nslots++;
                }
            }
            s.pass.ops = std::move(ops);
            s.pass.rt = true;
        }
        steps.push_back(std::move(s));
        rem.swap(rest);
    }
    return steps;
}

int low_bits(int dtype) { return dtype == PLK_C64 ? 4 : 3; }

int env_int(const char *name, int dflt) {
    const char *v = std::getenv(name);
    return v && *v ? std::atoi(v) : dflt;
}

// Tile width of backward (adjoint) passes; 0 selects register segments.
int tile_bits_adjoint(int n, int dtype) {
    if (env_int("PLK_TILES", 1) == 0) {
        return 0;
    }
    const int T = std::min({env_int("PLK_TILE_BITS_ADJ", 10), n - 8, TILE_TMAX});
    return T >= low_bits(dtype) + 2 ? T : 0;
}

int tile_bits_forward(int n, int dtype) {
    if (env_int("PLK_TILES", 1) == 0) {
        return 0;
    }
    const int T = std::min({env_int("PLK_TILE_BITS_FWD", dtype == PLK_C64 ? 11 : 10), n - 8, TILE_TMAX});
    return T >= low_bits(dtype) + 2 ? T : 0;
}

// Largest observable chunk whose tiles fit the LDS budget.
int tile_max_nj(int dtype, int T) {
    int best = 1;
    const int cap = env_int("PLK_TILE_NJ", 4);
    for (int nj = 1; nj <= 8; nj++) {
        if (tile_nj_supported(dtype, nj) &&
            tile_lds_bytes(dtype, T, nj, TILE_PMAX) <= TILE_LDS_MAX && nj <= cap) {
            best = nj;
        }
    }
    return best;
}

// n_obs split into balanced chunks of at most njmax.
std::vector<int> chunk_sizes(size_t n_obs, int njmax) {
    const size_t k = (n_obs + njmax - 1) / njmax;
    std::vector<int> out;
    for (size_t c = 0; c < k; c++) {
        out.push_back(static_cast<int>(n_obs / k + (c < n_obs % k ? 1 : 0)));
    }
    return out;
}

struct SegMeta {
    SegLaunch L{};
    size_t first = 0; // byte offset of the segment's op group in the table
    int cols[SEG_PMAX] = {0};
    double coefs[SEG_PMAX] = {0};
    TileLaunch TL{};
    size_t first_bundle = 0;
};

template <typename T>
void put_mat(T *dst, const std::vector<cd> &M, bool swap_bits) {
    const size_t dim = M.size() == 16 ? 4 : 2;
    auto sw = [&](size_t r) { return swap_bits ? (((r & 1) << 1) | ((r >> 1) & 1)) : r; };
    for (size_t r = 0; r < dim; r++) {
        for (size_t c = 0; c < dim; c++) {
            const cd v = M[sw(r) * dim + sw(c)];
            dst[2 * (r * dim + c)] = static_cast<T>(v.real());
            dst[2 * (r * dim + c) + 1] = static_cast<T>(v.imag());
        }
    }
}

template <typename T>
size_t pack_ops(std::vector<unsigned char> &out, const SegOp<T> *ops, size_t nops) {
    auto align16 = [&]() { out.resize((out.size() + 15) & ~size_t{15}, 0); };
    align16();
    const size_t base = out.size();
    out.resize(base + nops * sizeof(OpHead), 0);
    for (size_t o = 0; o < nops; o++) {
        const SegOp<T> &d = ops[o];
        const int dim = d.m == 1 ? 2 : 4;
        T u[32], g[32];
        int nu = 0, ng = 0;
        switch (d.ukind) {
        case U_DIAG:
            for (int r = 0; r < dim; r++) {
                u[nu++] = d.U[2 * (r * dim + r)];
                u[nu++] = d.U[2 * (r * dim + r) + 1];
            }
            break;
        case U_REAL:
            for (int e = 0; e < dim * dim; e++) {
                u[nu++] = d.U[2 * e];
            }
            break;
        case U_SWAP:
            break;
        default:
            for (int e = 0; e < 2 * dim * dim; e++) {
                u[nu++] = d.U[e];
            }
        }
        if (d.slot >= 0) {
            if (d.gkind == G_DIAG) {
                for (int r = 0; r < dim; r++) {
                    g[ng++] = d.G[2 * (r * dim + r)];
                }
            } else if (d.gkind == G_DENSE) {
                for (int e = 0; e < 2 * dim * dim; e++) {
                    g[ng++] = d.G[e];
                }
            }
        }
        align16();
        const size_t uoff = out.size() - base;
        out.insert(out.end(), reinterpret_cast<const unsigned char *>(u),
                   reinterpret_cast<const unsigned char *>(u + nu));
        align16();
        const size_t goff = out.size() - base;
        out.insert(out.end(), reinterpret_cast<const unsigned char *>(g),
                   reinterpret_cast<const unsigned char *>(g + ng));
        if (goff > 0xffff) {
            raise(PLK_ERR_UNSUPPORTED, "operation group exceeds 64 KiB");
        }
        OpHead h{};
        h.cls = d.cls;
        h.slot = d.slot;
        h.kinds = d.ukind | (d.gkind << 8) | (d.swap << 16);
        h.offs = static_cast<int32_t>(uoff | (goff << 16));
        std::memcpy(out.data() + base + o * sizeof(OpHead), &h, sizeof(OpHead));
    }
    align16();
    return base;
}

// Picks the structure-specialized kernel path for a descriptor whose U (and
// G) already hold the final, bit-permuted matrices of the op.
template <typename T> void classify(SegOp<T> &d, size_t m) {
    const int dim = m == 1 ? 2 : 4;
    auto re = [&](const T *M, int r, int c) { return M[2 * (r * dim + c)]; };
    auto im = [&](const T *M, int r, int c) { return M[2 * (r * dim + c) + 1]; };
    bool diag = true, real = true;
    for (int r = 0; r < dim; r++) {
        for (int c = 0; c < dim; c++) {
            if (r != c && (re(d.U, r, c) != 0 || im(d.U, r, c) != 0)) {
                diag = false;
            }
            if (im(d.U, r, c) != 0) {
                real = false;
            }
        }
    }
    // A transposition of two op-local indices with unit entries.
    int moved[4], nm = 0;
    bool perm = real;
    for (int r = 0; r < dim && perm; r++) {
        int ones = 0, col = -1;
        for (int c = 0; c < dim; c++) {
            const T v = re(d.U, r, c);
            if (v == 1) {
                ones++;
                col = c;
            } else if (v != 0) {
                perm = false;
            }
        }
        perm = perm && ones == 1;
        if (perm && col != r) {
            if (nm < 4) {
                moved[nm] = r;
            }
            nm++;
        }
    }
    if (diag) {
        d.ukind = U_DIAG;
    } else if (perm && nm == 2 && re(d.U, moved[0], moved[1]) == 1) {
        d.ukind = U_SWAP;
        d.swap = 4 * moved[0] + moved[1];
    } else if (real) {
        d.ukind = U_REAL;
    } else {
        d.ukind = U_DENSE;
    }
    d.gkind = G_DENSE;
    if (d.slot >= 0) {
        bool gdiag = true;
        for (int r = 0; r < dim; r++) {
            for (int c = 0; c < dim; c++) {
                if (im(d.G, r, c) != 0 || (r != c && re(d.G, r, c) != 0)) {
                    gdiag = false;
                }
            }
        }
        bool gz = gdiag;
        for (int r = 0; r < dim && gz; r++) {
            gz = re(d.G, r, r) == ((__builtin_popcount(r) & 1) ? -1 : 1);
        }
        if (gz) {
            d.gkind = G_Z;
        } else if (gdiag) {
            d.gkind = G_DIAG;
        } else if (m == 1 && re(d.G, 0, 0) == 0 && im(d.G, 0, 0) == 0 && re(d.G, 1, 1) == 0 &&
                   im(d.G, 1, 1) == 0) {
            if (re(d.G, 0, 1) == 1 && im(d.G, 0, 1) == 0 && re(d.G, 1, 0) == 1 && im(d.G, 1, 0) == 0) {
                d.gkind = G_X;
            } else if (re(d.G, 0, 1) == 0 && im(d.G, 0, 1) == -1 && re(d.G, 1, 0) == 0 &&
                       im(d.G, 1, 0) == 1) {
                d.gkind = G_Y;
            }
        }
    }
}

template <typename T>
SegMeta build_segment(const Segment &seg, int n, std::vector<unsigned char> &table) {
    SegMeta meta;
    SegLaunch &L = meta.L;
    L.n = n;
    L.K = static_cast<int>(seg.wires.size());
    std::vector<std::pair<int, size_t>> bw; // (bit, wire)
    for (auto w : seg.wires) {
        bw.push_back({n - 1 - static_cast<int>(w), w});
    }
    std::sort(bw.begin(), bw.end());
    std::map<size_t, int> pos;
    for (int q = 0; q < L.K; q++) {
        L.bits[q] = bw[q].first;
        pos[bw[q].second] = q;
    }
    std::vector<SegOp<T>> group;
    int slot = 0;
    for (const auto &p : seg.ops) {
        SegOp<T> d{};
        d.slot = -1;
        d.m = static_cast<int32_t>(p.U.wires.size());
        const auto &w = p.U.wires;
        bool swapb = false;
        if (w.size() == 1) {
            d.cls = pos.at(w[0]);
        } else {
            int qa = pos.at(w[0]), qb = pos.at(w[1]);
            if (qa < qb) {
                std::swap(qa, qb);
                swapb = true;
            }
            if (L.K == 2) {
                d.cls = 2;
            } else {
                d.cls = (qa == 1) ? 3 : (qb == 0 ? 4 : 5);
            }
        }
        put_mat(d.U, p.U.mat, swapb);
        if (p.trainable) {
            if (p.G.wires != p.U.wires) {
                raise(PLK_ERR_INVALID, "generator wires differ from gate wires");
            }
            put_mat(d.G, p.G.mat, swapb);
            d.slot = slot;
            meta.cols[slot] = p.col;
            meta.coefs[slot] = p.coef;
            slot++;
        }
        classify(d, w.size());
        group.push_back(d);
    }
    meta.first = pack_ops(table, group.data(), group.size());
    L.nops = static_cast<int>(seg.ops.size());
    L.nslots = slot;
    return meta;
}

// U payload layout a patched gate keeps for every angle (the structure of
// the gate at a generic angle, which special angles only refine).
int32_t patch_ukind(int16_t gate) {
    switch (gate) {
    case BPG_RX:
        return U_DENSE;
    case BPG_RY:
        return U_REAL;
    default:
        return U_DIAG;
    }
}

template <typename T>
void build_pass(const Pass &ps, int n, std::vector<unsigned char> &table,
                std::vector<TileBundle> &btab, SegMeta &meta,
                std::vector<BatchPatch> *patches = nullptr) {
    TileLaunch &TL = meta.TL;
    TL.n = n;
    TL.T = static_cast<int>(ps.tb.size());
    for (int q = 0; q < TL.T; q++) {
        TL.tb[q] = ps.tb[q];
    }
    int pos[64];
    for (size_t q = 0; q < ps.tb.size(); q++) {
        pos[ps.tb[q]] = static_cast<int>(q);
    }
    meta.first_bundle = btab.size();
    int slot0 = 0;
    for (const auto &b : ps.bundles) {
        TileBundle tb{};
        tb.K = b.K;
        tb.qa = b.qa;
        tb.qb = b.qb;
        tb.nops = static_cast<int32_t>(b.ops.size());
        tb.nslots = b.nslots;
        tb.slot0 = slot0;
        int slot = 0;
        std::vector<SegOp<T>> group;
        for (const auto &p : b.ops) {
            SegOp<T> d{};
            d.slot = -1;
            d.m = static_cast<int32_t>(p.U.wires.size());
            const auto &w = p.U.wires;
            bool swapb = false;
            if (w.size() == 1) {
                const int q = pos[n - 1 - static_cast<int>(w[0])];
                d.cls = (b.K == 1 || q == b.qb) ? 0 : 1;
            } else {
                const int q0 = pos[n - 1 - static_cast<int>(w[0])];
                swapb = (q0 != b.qa);
                d.cls = 2;
            }
            put_mat(d.U, p.U.mat, swapb);
            if (p.trainable) {
                put_mat(d.G, p.G.mat, swapb);
                d.slot = slot++;
            }
            classify(d, w.size());
            if (patches != nullptr && p.patch_gate != BPG_NONE) {
                d.ukind = patch_ukind(p.patch_gate);
            }
            group.push_back(d);
        }
        const size_t off = pack_ops(table, group.data(), group.size());
        if (off > static_cast<size_t>(INT32_MAX)) {
            raise(PLK_ERR_UNSUPPORTED, "operation table exceeds 2 GiB");
        }
        if (patches != nullptr) {
            for (size_t o = 0; o < b.ops.size(); o++) {
                const PlanOp &p = b.ops[o];
                if (p.patch_gate == BPG_NONE) {
                    continue;
                }
                OpHead h;
                std::memcpy(&h, table.data() + off + o * sizeof(OpHead), sizeof(OpHead));
                patches->push_back(BatchPatch{static_cast<int64_t>(off) + (h.offs & 0xffff),
                                              p.param, p.patch_gate,
                                              static_cast<int16_t>(p.adj ? 1 : 0)});
            }
        }
        tb.first = static_cast<int32_t>(off);
        slot0 += b.nslots;
        btab.push_back(tb);
    }
    TL.nbundles = static_cast<int>(ps.bundles.size());
    TL.nslots = ps.nslots;
}

// Device tables of segment/pass descriptors for a whole plan (one upload).
// Host images of a plan's device tables: op groups of register segments and
// LDS bundles, the bundle table, and register-tile programs, with the
// launch metadata that points into them.
struct RtProgram {
    RtLaunch L{};
    size_t off = 0; // byte offset of the program in the plan blob
};

// A single-op step run by the step kernels (kernels_step.hip): its launch
// arguments, the byte offset of its U and G matrices in the image's matrix
// table, and its Jacobian column and coefficient.
struct StepMeta {
    bool ok = false;
    StepLaunch L{};
    size_t off = 0;
    int col = 0;
    double coef = 0.0;
};

struct PlanImage {
    std::vector<unsigned char> ops;
    std::vector<TileBundle> bundles;
    std::vector<SegMeta> metas;
    std::vector<unsigned char> rt;
    std::vector<RtProgram> progs;
    std::vector<unsigned char> mats; // U and G of the step-kernel steps (FP64 complex)
    std::vector<StepMeta> singles;
};

// Device addresses of an image's tables.
struct PlanPlace {
    const unsigned char *ops = nullptr;
    const TileBundle *bundles = nullptr;
    const unsigned char *rt = nullptr;
    const unsigned char *mats = nullptr;
};

// The 2^m x 2^m matrix, over the wires uw (uw[0] most significant), of an
// operator on the wires gw (a subset of uw) given as a dense matrix or as
// parity phases (even / odd); false when gw is not a subset of uw.
bool embed_matrix(const std::vector<size_t> &uw, const std::vector<size_t> &gw, OpKind kind,
                  const std::vector<cd> &gm, cd even, cd odd, std::vector<cd> &out) {
    const int m = static_cast<int>(uw.size());
    const int k = static_cast<int>(gw.size());
    const int D = 1 << m;
    std::vector<int> shift; // bit of the local index that carries gw[q]
    int gmask = 0;
    for (auto w : gw) {
        const auto it = std::find(uw.begin(), uw.end(), w);
        if (it == uw.end()) {
            return false;
        }
        shift.push_back(m - 1 - static_cast<int>(it - uw.begin()));
        gmask |= 1 << shift.back();
    }
    auto sub = [&](int r) {
        int v = 0;
        for (int q = 0; q < k; q++) {
            v |= ((r >> shift[q]) & 1) << (k - 1 - q);
        }
        return v;
    };
    out.assign(static_cast<size_t>(D) * D, cd{0.0, 0.0});
    for (int r = 0; r < D; r++) {
        for (int c = 0; c < D; c++) {
            if ((r & ~gmask) != (c & ~gmask)) {
                continue;
            }
            const int sr = sub(r), sc = sub(c);
            cd v{0.0, 0.0};
            if (kind == OpKind::Dense) {
                v = gm[static_cast<size_t>(sr) * (size_t{1} << k) + sc];
            } else if (sr == sc) {
                v = (__builtin_popcount(sr) & 1) ? odd : even;
            }
            out[static_cast<size_t>(r) * D + c] = v;
        }
    }
    return true;
}

// Fills m for a single-op step the step kernels can run: parity phases on
// any number of wires (with a parity generator when trainable) or a dense
// matrix on at most 4 wires whose generator acts inside its wires, restricted
// to the local states it moves. Appends the matrices to mats; returns false
// for anything else (the op then runs through the per-state path).
bool build_step(const PlanOp &p, int n, std::vector<unsigned char> &mats, StepMeta &m) {
    if (env_int("PLK_STEPS", 1) == 0) {
        return false;
    }
    const CompiledOp &U = p.U;
    StepLaunch L{};
    auto mask = [&](const std::vector<size_t> &w) {
        uint64_t x = 0;
        for (auto v : w) {
            x |= uint64_t{1} << (n - 1 - static_cast<int>(v));
        }
        return x;
    };
    if (U.kind == OpKind::Parity) {
        if (p.trainable && p.G.kind != OpKind::Parity) {
            return false;
        }
        L.m = 0;
        L.groups = uint64_t{1} << n;
        L.umask = mask(U.wires);
        L.gmask = p.trainable ? mask(p.G.wires) : 0;
        L.ue[0] = U.even.real();
        L.ue[1] = U.even.imag();
        L.uo[0] = U.odd.real();
        L.uo[1] = U.odd.imag();
    } else if (U.kind == OpKind::Dense && !U.wires.empty() && U.wires.size() <= STEP_MMAX) {
        const int mw = static_cast<int>(U.wires.size());
        const size_t full = size_t{1} << mw;
        std::vector<cd> gm(full * full, cd{0.0, 0.0});
        if (p.trainable &&
            !embed_matrix(U.wires, p.G.wires, p.G.kind, p.G.mat, cd{1.0, 0.0}, cd{-1.0, 0.0}, gm)) {
            return false;
        }
        // The local states the step has to touch: those where U differs
        // from the identity or G is non-zero (exact comparisons: any
        // deviation keeps a state in), padded to 2, 4, 8 or 16 with states
        // on which U is the identity and G zero.
        std::vector<size_t> act, idle;
        for (size_t k = 0; k < full; k++) {
            bool on = false;
            for (size_t c = 0; c < full && !on; c++) {
                const cd e = k == c ? cd{1.0, 0.0} : cd{0.0, 0.0};
                on = U.mat[k * full + c] != e || U.mat[c * full + k] != e ||
                     gm[k * full + c] != cd{0.0, 0.0} || gm[c * full + k] != cd{0.0, 0.0};
            }
            (on ? act : idle).push_back(k);
        }
        size_t D = 2;
        while (D < act.size()) {
            D *= 2;
        }
        for (size_t q = 0; act.size() < D; q++) {
            act.push_back(idle[q]);
        }
        L.m = mw;
        L.dim = static_cast<int>(D);
        L.groups = uint64_t{1} << (n - mw);
        std::vector<int> bits;
        for (auto w : U.wires) {
            bits.push_back(n - 1 - static_cast<int>(w));
        }
        std::vector<int> sorted = bits;
        std::sort(sorted.begin(), sorted.end());
        for (int q = 0; q < mw; q++) {
            L.bits[q] = sorted[q];
        }
        for (size_t i = 0; i < D; i++) {
            uint64_t o = 0;
            for (int q = 0; q < mw; q++) {
                if ((act[i] >> (mw - 1 - q)) & 1U) {
                    o |= uint64_t{1} << bits[q];
                }
            }
            L.off[i] = o;
        }
        mats.resize((mats.size() + 255) & ~size_t{255}, 0);
        m.off = mats.size();
        auto put = [&](const std::vector<cd> &M) {
            for (size_t r = 0; r < D; r++) {
                for (size_t c = 0; c < D; c++) {
                    const cd v = M[act[r] * full + act[c]];
                    const double re = v.real(), im = v.imag();
                    const auto *a = reinterpret_cast<const unsigned char *>(&re);
                    const auto *b = reinterpret_cast<const unsigned char *>(&im);
                    mats.insert(mats.end(), a, a + sizeof re);
                    mats.insert(mats.end(), b, b + sizeof im);
                }
            }
        };
        put(U.mat);
        put(gm);
    } else {
        return false;
    }
    L.trainable = p.trainable ? 1 : 0;
    L.write_phi = 1;
    m.L = L;
    m.col = p.col;
    m.coef = p.coef;
    m.ok = true;
    return true;
}

// Whether the step kernels run this single-op step.
bool step_kernel_ok(const PlanOp &p, int n) {
    std::vector<unsigned char> scratch;
    StepMeta m;
    return build_step(p, n, scratch, m);
}

template <typename T>
void build_seg_tables(const std::vector<Step> &steps, int n, PlanImage &img) {
    for (const auto &s : steps) {
        if (s.fused) {
            img.metas.push_back(build_segment<T>(s.seg, n, img.ops));
        } else if (s.tiled && !s.pass.rt) {
            SegMeta m;
            build_pass<T>(s.pass, n, img.ops, img.bundles, m);
            img.metas.push_back(m);
        } else {
            img.metas.push_back(SegMeta{});
        }
    }
}

// ---------------------------------------------------------------------------
// Register-tile pass programs (kernels_rtile.hip).
//
// A pass's tile bits are split into R register bits and 6 lane positions.
// An operation needs in registers the bits it mixes (bits in which its U, or
// its generator, has an entry between different values); its other bits may
// sit on lanes, where each lane applies the block its lane bits select. The
// builder picks the load layout (the first mixing bits in registers, the
// next two on lane positions 5 and 4, whose exchanges are single
// v_permlane swaps), then walks the ops and brings each missing mixing bit
// into a register by exchanging it with the register whose next use is
// farthest away (Belady). The final layout is the store layout.
// ---------------------------------------------------------------------------
int rt_register_bits(int dtype) { return dtype == PLK_C64 ? 4 : 3; }

template <typename T> struct RtBuilder {
    std::vector<unsigned char> &blob;
    size_t base = 0;
    std::vector<RtInstr> ins;
    std::vector<unsigned char> pay;

    explicit RtBuilder(std::vector<unsigned char> &b) : blob(b) {}

    int32_t put(const std::vector<double> &vals) {
        pay.resize((pay.size() + 15) & ~size_t{15}, 0);
        const size_t off = pay.size();
        for (double v : vals) {
            const T t = static_cast<T>(v);
            const auto *p = reinterpret_cast<const unsigned char *>(&t);
            pay.insert(pay.end(), p, p + sizeof(T));
        }
        return static_cast<int32_t>(off); // relative to the payload start; rebased in finish()
    }

    size_t finish() {
        const size_t head = ins.size() * sizeof(RtInstr);
        for (auto &i : ins) {
            i.uoff += static_cast<int32_t>(head);
            i.goff += static_cast<int32_t>(head);
        }
        blob.resize((blob.size() + 255) & ~size_t{255}, 0);
        const size_t off = blob.size();
        if (off + head + pay.size() > static_cast<size_t>(INT32_MAX)) {
            raise(PLK_ERR_UNSUPPORTED, "register-tile program table exceeds 2 GiB");
        }
        const auto *pi = reinterpret_cast<const unsigned char *>(ins.data());
        blob.insert(blob.end(), pi, pi + head);
        blob.insert(blob.end(), pay.begin(), pay.end());
        return off;
    }
};

std::vector<double> cplx(const std::vector<cd> &v) {
    std::vector<double> out;
    for (auto &x : v) {
        out.push_back(x.real());
        out.push_back(x.imag());
    }
    return out;
}

bool is_real(const std::vector<cd> &v) {
    for (auto &x : v) {
        if (x.imag() != 0.0) {
            return false;
        }
    }
    return true;
}

// 2x2 block of a 4x4 op matrix over local bit `lr` (the register wire) with
// the other local bit fixed to c.
std::vector<cd> block2(const std::vector<cd> &M, int lr, int c) {
    std::vector<cd> B(4);
    for (int x = 0; x < 2; x++) {
        for (int y = 0; y < 2; y++) {
            const int r = lr == 1 ? (2 * x + c) : (2 * c + x);
            const int col = lr == 1 ? (2 * y + c) : (2 * c + y);
            B[2 * x + y] = M[4 * r + col];
        }
    }
    return B;
}

bool is_diag2(const std::vector<cd> &B) { return B[1] == cd(0, 0) && B[2] == cd(0, 0); }
bool is_ident2(const std::vector<cd> &B) {
    return B[0] == cd(1, 0) && B[3] == cd(1, 0) && is_diag2(B);
}
bool is_x2(const std::vector<cd> &B) {
    return B[0] == cd(0, 0) && B[3] == cd(0, 0) && B[1] == cd(1, 0) && B[2] == cd(1, 0);
}

// Encodes the U (and, for a trainable op, G) parts of a one-register-bit op.
template <typename T>
void rt_encode_reg1(RtBuilder<T> &b, RtInstr &in, const std::vector<cd> &U, const std::vector<cd> *G) {
    if (is_diag2(U)) {
        in.ukind = U_DIAG;
        in.uoff = b.put(cplx({U[0], U[3]}));
    } else if (is_x2(U)) {
        in.ukind = U_SWAP;
    } else if (is_real(U)) {
        in.ukind = U_REAL;
        in.uoff = b.put({U[0].real(), U[1].real(), U[2].real(), U[3].real()});
    } else {
        in.ukind = U_DENSE;
        in.uoff = b.put(cplx(U));
    }
    if (G) {
        const auto &g = *G;
        if (is_diag2(g) && is_real(g) && g[0] == cd(1, 0) && g[3] == cd(-1, 0)) {
            in.gkind = G_Z;
        } else if (is_diag2(g) && is_real(g)) {
            in.gkind = G_DIAG;
            in.goff = b.put({g[0].real(), g[3].real()});
        } else if (is_x2(g)) {
            in.gkind = G_X;
        } else if (g[0] == cd(0, 0) && g[3] == cd(0, 0) && g[1] == cd(0, -1) && g[2] == cd(0, 1)) {
            in.gkind = G_Y;
        } else {
            in.gkind = G_DENSE;
            in.goff = b.put(cplx(g));
        }
    }
}

template <typename T>
RtProgram build_rt_program(const Pass &ps, int n, int R, std::vector<unsigned char> &blob) {
    RtProgram out;
    RtLaunch &L = out.L;
    L.n = n;
    L.R = R;
    L.T = static_cast<int>(ps.tb.size());
    if (L.T != R + RT_LANE_BITS) {
        raise(PLK_ERR_INVALID, "register-tile pass of the wrong width");
    }
    for (int q = 0; q < L.T; q++) {
        L.tb[q] = ps.tb[q];
    }
    const size_t nops = ps.ops.size();
    std::vector<std::vector<int>> gbits(nops), mix(nops);
    for (size_t i = 0; i < nops; i++) {
        const PlanOp &p = ps.ops[i];
        const int m = static_cast<int>(p.U.wires.size());
        const int dim = 1 << m;
        for (int k = 0; k < m; k++) {
            const int gb = n - 1 - static_cast<int>(p.U.wires[k]);
            gbits[i].push_back(gb);
            const int lb = m - 1 - k;
            if (mixes_bit(p.U.mat, dim, lb) || (p.trainable && mixes_bit(p.G.mat, dim, lb))) {
                mix[i].push_back(gb);
            }
        }
    }
    // Load layout.
    std::vector<int> order;
    for (size_t i = 0; i < nops; i++) {
        for (int q : mix[i]) {
            if (std::find(order.begin(), order.end(), q) == order.end()) {
                order.push_back(q);
            }
        }
    }
    // The lowest `keep` index bits sit on lanes at load and store (complex128:
    // bit 0, complex64: bits 0 and 1), so every load and store instruction
    // covers runs of at least 32 bytes. With the three lowest bits all in
    // registers each instruction touches 64 cache lines 16 bytes at a time,
    // and an in-place pass streams at 4.4 TB/s instead of 5.1 (MI355X,
    // 1-4 GiB states). A keep bit the pass mixes loads on lane position 5
    // or 4 and is swapped into a register when an op needs it.
    const int keep = std::is_same<T, double>::value ? 1 : 2;
    auto is_keep = [&](int q) { return q < keep; };
    auto has = [](const std::vector<int> &v, int q) { return std::find(v.begin(), v.end(), q) != v.end(); };
    std::vector<int> reg, lane(RT_LANE_BITS, -1);
    std::vector<int> rest;
    for (size_t k = 0; k < order.size() && static_cast<int>(reg.size()) < R; k++) {
        if (!is_keep(order[k])) {
            reg.push_back(order[k]);
        }
    }
    std::vector<int> hot; // mixing bits not in registers: lane positions 5, 4
    for (int q : order) {
        if (is_keep(q) && hot.size() < 2) {
            hot.push_back(q);
        }
    }
    for (size_t k = 0; k < order.size() && hot.size() < 2; k++) {
        if (!is_keep(order[k]) && !has(reg, order[k])) {
            hot.push_back(order[k]);
        }
    }
    for (int q : ps.tb) {
        if (!has(reg, q) && !has(hot, q)) {
            rest.push_back(q); // ascending global bits
        }
    }
    while (static_cast<int>(reg.size()) < R) {
        // Fill registers with the highest remaining bits, never a keep bit.
        reg.push_back(rest.back());
        rest.pop_back();
    }
    for (size_t k = 0; k < hot.size(); k++) {
        lane[5 - k] = hot[k];
    }
    for (int p = 0, r = 0; p < RT_LANE_BITS; p++) {
        if (lane[p] < 0) {
            lane[p] = rest[r++];
        }
    }
    for (int r = 0; r < R; r++) {
        L.reg_in[r] = reg[r];
    }
    for (int p = 0; p < RT_LANE_BITS; p++) {
        L.lane_in[p] = lane[p];
    }

    RtBuilder<T> b(blob);
    auto reg_of = [&](int q) {
        for (int r = 0; r < R; r++) {
            if (reg[r] == q) {
                return r;
            }
        }
        return -1;
    };
    auto lane_of = [&](int q) {
        for (int p = 0; p < RT_LANE_BITS; p++) {
            if (lane[p] == q) {
                return p;
            }
        }
        raise(PLK_ERR_INVALID, "register-tile op outside the tile bits");
    };
    // Selector of a non-mixing bit: its lane position, or 64 + the bit when
    // it lies outside the tile (fixed for the whole tile).
    auto sel_of = [&](int q) -> int {
        for (int p = 0; p < RT_LANE_BITS; p++) {
            if (lane[p] == q) {
                return p;
            }
        }
        if (std::find(ps.tb.begin(), ps.tb.end(), q) != ps.tb.end()) {
            raise(PLK_ERR_INVALID, "register-tile selector bit sits in a register");
        }
        return 64 + q;
    };
    auto next_use = [&](int q, size_t from) {
        for (size_t j = from; j < nops; j++) {
            if (std::find(mix[j].begin(), mix[j].end(), q) != mix[j].end()) {
                return j;
            }
        }
        return nops + 1;
    };
    int slot = 0;
    for (size_t i = 0; i < nops; i++) {
        const PlanOp &p = ps.ops[i];
        RtInstr in{};
        in.slot = -1;
        in.s0 = in.s1 = -1;
        for (int q : mix[i]) {
            if (reg_of(q) >= 0) {
                continue;
            }
            int victim = -1;
            size_t far = 0;
            for (int r = 0; r < R; r++) {
                if (std::find(mix[i].begin(), mix[i].end(), reg[r]) != mix[i].end()) {
                    continue;
                }
                const size_t u = next_use(reg[r], i);
                if (victim < 0 || u > far) {
                    victim = r;
                    far = u;
                }
            }
            const int l = lane_of(q);
            if (l < 4 || in.nremap >= 2) {
                raise(PLK_ERR_INVALID, "register-tile pass needs more mixing bits than it can remap");
            }
            if (in.nremap == 0) {
                in.rq0 = static_cast<int8_t>(victim);
                in.rp0 = static_cast<int8_t>(l);
            } else {
                in.rq1 = static_cast<int8_t>(victim);
                in.rp1 = static_cast<int8_t>(l);
            }
            in.nremap++;
            lane[l] = reg[victim];
            reg[victim] = q;
        }
        const std::vector<cd> *G = nullptr;
        if (p.trainable) {
            if (p.G.wires != p.U.wires || p.G.kind != OpKind::Dense) {
                raise(PLK_ERR_INVALID, "generator wires differ from gate wires");
            }
            G = &p.G.mat;
            in.slot = slot++;
        }
        const int m = static_cast<int>(gbits[i].size());
        if (m == 1) {
            const int r = reg_of(gbits[i][0]);
            if (r >= 0) {
                in.code = RT_REG1;
                in.a = static_cast<int8_t>(r);
                rt_encode_reg1<T>(b, in, p.U.mat, G);
            } else {
                in.code = RT_LANE;
                in.s0 = static_cast<int8_t>(sel_of(gbits[i][0]));
                in.uoff = b.put(cplx({p.U.mat[0], p.U.mat[3]}));
                if (G) {
                    in.goff = b.put({(*G)[0].real(), (*G)[3].real()});
                }
            }
        } else {
            const int r0 = reg_of(gbits[i][0]), r1 = reg_of(gbits[i][1]);
            if (r0 >= 0 && r1 >= 0) {
                in.code = RT_REG2;
                in.a = static_cast<int8_t>(std::max(r0, r1));
                in.b = static_cast<int8_t>(std::min(r0, r1));
                SegOp<T> d{};
                d.slot = G ? 0 : -1;
                put_mat(d.U, p.U.mat, r0 < r1);
                if (G) {
                    put_mat(d.G, *G, r0 < r1);
                }
                classify(d, 2);
                in.ukind = static_cast<int8_t>(d.ukind);
                in.swap = static_cast<int8_t>(d.swap);
                std::vector<double> u;
                if (d.ukind == U_DIAG) {
                    for (int k = 0; k < 4; k++) {
                        u.push_back(d.U[2 * (5 * k)]);
                        u.push_back(d.U[2 * (5 * k) + 1]);
                    }
                } else if (d.ukind == U_REAL) {
                    for (int e = 0; e < 16; e++) {
                        u.push_back(d.U[2 * e]);
                    }
                } else if (d.ukind == U_DENSE) {
                    u.assign(d.U, d.U + 32);
                }
                in.uoff = b.put(u);
                if (G) {
                    in.gkind = static_cast<int8_t>(d.gkind);
                    std::vector<double> g;
                    if (d.gkind == G_DIAG) {
                        for (int k = 0; k < 4; k++) {
                            g.push_back(d.G[2 * (5 * k)]);
                        }
                    } else if (d.gkind == G_DENSE) {
                        g.assign(d.G, d.G + 32);
                    }
                    in.goff = b.put(g);
                }
            } else if (r0 >= 0 || r1 >= 0) {
                // One wire on a register bit, the other (not mixed) on a lane.
                const int k_reg = r0 >= 0 ? 0 : 1;
                const int lr = 1 - k_reg; // local bit of the register wire
                in.code = RT_SEL;
                in.a = static_cast<int8_t>(k_reg == 0 ? r0 : r1);
                in.s0 = static_cast<int8_t>(sel_of(gbits[i][1 - k_reg]));
                const std::vector<cd> B0 = block2(p.U.mat, lr, 0), B1 = block2(p.U.mat, lr, 1);
                if (is_diag2(B0) && is_diag2(B1)) {
                    in.ukind = U_DIAG;
                    in.uoff = b.put(cplx({B0[0], B0[3], B1[0], B1[3]}));
                } else if ((is_ident2(B0) || is_x2(B0)) && (is_ident2(B1) || is_x2(B1))) {
                    in.ukind = U_SWAP;
                    in.uoff = b.put({is_x2(B0) ? 1.0 : 0.0, is_x2(B1) ? 1.0 : 0.0});
                } else if (is_real(B0) && is_real(B1)) {
                    in.ukind = U_REAL;
                    in.uoff = b.put({B0[0].real(), B0[1].real(), B0[2].real(), B0[3].real(),
                                     B1[0].real(), B1[1].real(), B1[2].real(), B1[3].real()});
                } else {
                    in.ukind = U_DENSE;
                    std::vector<double> u = cplx(B0), u1 = cplx(B1);
                    u.insert(u.end(), u1.begin(), u1.end());
                    in.uoff = b.put(u);
                }
                if (G) {
                    const std::vector<cd> G0 = block2(*G, lr, 0), G1 = block2(*G, lr, 1);
                    if (is_diag2(G0) && is_diag2(G1) && is_real(G0) && is_real(G1)) {
                        in.gkind = G_DIAG;
                        in.goff = b.put({G0[0].real(), G0[3].real(), G1[0].real(), G1[3].real()});
                    } else {
                        in.gkind = G_DENSE;
                        std::vector<double> g = cplx(G0), g1 = cplx(G1);
                        g.insert(g.end(), g1.begin(), g1.end());
                        in.goff = b.put(g);
                    }
                }
            } else {
                // Both wires on lanes: diagonal; block c = bit(wires[1]) + 2 bit(wires[0]) is
                // the op's local index.
                in.code = RT_LANE;
                in.s0 = static_cast<int8_t>(sel_of(gbits[i][1]));
                in.s1 = static_cast<int8_t>(sel_of(gbits[i][0]));
                std::vector<cd> f;
                for (int k = 0; k < 4; k++) {
                    f.push_back(p.U.mat[5 * k]);
                }
                in.uoff = b.put(cplx(f));
                if (G) {
                    std::vector<double> g;
                    for (int k = 0; k < 4; k++) {
                        g.push_back((*G)[5 * k].real());
                    }
                    in.goff = b.put(g);
                }
            }
        }
        b.ins.push_back(in);
    }
    // Store layout: keep bits left in registers go back to lane positions 5
    // or 4 (whichever holds a bit that is not a keep bit).
    L.nfin = 0;
    for (int q = 0; q < keep; q++) {
        const int r = reg_of(q);
        if (r < 0) {
            continue;
        }
        const int p = !is_keep(lane[5]) ? 5 : 4;
        if (is_keep(lane[p])) {
            raise(PLK_ERR_INVALID, "register-tile store layout: no lane position for a low bit");
        }
        L.fin_rq[L.nfin] = r;
        L.fin_rp[L.nfin] = p;
        L.nfin++;
        std::swap(reg[r], lane[p]);
    }
    for (int r = 0; r < R; r++) {
        L.reg_out[r] = reg[r];
    }
    for (int p = 0; p < RT_LANE_BITS; p++) {
        L.lane_out[p] = lane[p];
    }
    L.ninstr = static_cast<int>(b.ins.size());
    L.nslots = slot;
    out.off = b.finish();
    return out;
}

// Programs of every register-tile pass of a plan, uploaded as one blob.
PlanImage build_image(const std::vector<Step> &steps, const State &s) {
    PlanImage img;
    if (s.dtype == PLK_C64) {
        build_seg_tables<float>(steps, s.n, img);
    } else {
        build_seg_tables<double>(steps, s.n, img);
    }
    img.singles.resize(steps.size());
    for (size_t i = 0; i < steps.size(); i++) {
        if (!steps[i].fused && !steps[i].tiled) {
            build_step(steps[i].single, s.n, img.mats, img.singles[i]);
        }
    }
    img.progs.resize(steps.size());
    const int R = rt_register_bits(s.dtype);
    for (size_t i = 0; i < steps.size(); i++) {
        if (steps[i].tiled && steps[i].pass.rt) {
            img.progs[i] = s.dtype == PLK_C64 ? build_rt_program<float>(steps[i].pass, s.n, R, img.rt)
                                              : build_rt_program<double>(steps[i].pass, s.n, R, img.rt);
        }
    }
    return img;
}

size_t image_bytes(const PlanImage &img) {
    auto up = [](size_t x) { return (x + 255) & ~size_t{255}; };
    return up(img.ops.size()) + up(img.bundles.size() * sizeof(TileBundle)) + up(img.rt.size()) +
           up(img.mats.size());
}

// Copies the image's tables to base (image_bytes(img) bytes) on the device's
// stream; the image must stay alive until the stream has run the copies.
PlanPlace copy_image(Device &dev, const PlanImage &img, unsigned char *base) {
    auto up = [](size_t x) { return (x + 255) & ~size_t{255}; };
    PlanPlace pl;
    size_t off = 0;
    auto put = [&](const void *src, size_t bytes) -> unsigned char * {
        unsigned char *dst = base + off;
        if (bytes) {
            PLK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, dev.stream));
        }
        off += up(bytes);
        return dst;
    };
    pl.ops = put(img.ops.data(), img.ops.size());
    pl.bundles = reinterpret_cast<const TileBundle *>(
        put(img.bundles.data(), img.bundles.size() * sizeof(TileBundle)));
    pl.rt = put(img.rt.data(), img.rt.size());
    pl.mats = put(img.mats.data(), img.mats.size());
    return pl;
}

// The addresses copy_image(dev, img, base) fills, without copying.
PlanPlace place_addresses(const PlanImage &img, unsigned char *base) {
    auto up = [](size_t x) { return (x + 255) & ~size_t{255}; };
    PlanPlace pl;
    pl.ops = base;
    pl.bundles = reinterpret_cast<const TileBundle *>(base + up(img.ops.size()));
    pl.rt = base + up(img.ops.size()) + up(img.bundles.size() * sizeof(TileBundle));
    pl.mats = pl.rt + up(img.rt.size());
    return pl;
}

PlanPlace place_in_arena(Device &dev, const PlanImage &img) {
    const size_t bytes = image_bytes(img);
    return copy_image(dev, img, static_cast<unsigned char *>(arena_alloc(dev, bytes ? bytes : 256)));
}

// Everything a replay of the plan's launches takes as kernel arguments:
// equal signatures mean equal launches, whatever the matrices in the tables.
void append_signature(const PlanImage &img, std::vector<unsigned char> &sig) {
    auto put = [&](const void *p, size_t bytes) {
        const auto *b = static_cast<const unsigned char *>(p);
        sig.insert(sig.end(), b, b + bytes);
    };
    const uint64_t sizes[4] = {img.ops.size(), img.bundles.size(), img.rt.size(), img.mats.size()};
    put(sizes, sizeof sizes);
    for (const StepMeta &m : img.singles) {
        put(&m.ok, sizeof m.ok);
        put(&m.L, sizeof m.L);
        put(&m.off, sizeof m.off);
        put(&m.col, sizeof m.col);
        put(&m.coef, sizeof m.coef);
    }
    for (const SegMeta &m : img.metas) {
        put(&m.L, sizeof m.L);
        put(&m.first, sizeof m.first);
        put(m.cols, sizeof m.cols);
        put(m.coefs, sizeof m.coefs);
        put(&m.TL, sizeof m.TL);
        put(&m.first_bundle, sizeof m.first_bundle);
    }
    for (const RtProgram &p : img.progs) {
        put(&p.L, sizeof p.L);
        put(&p.off, sizeof p.off);
    }
}

// Observable states per register-tile chunk: 3 keeps phi plus three states
// and the spare buffer at two wavefronts per SIMD. When a pass has fewer
// tiles than the device has CUs, chunks of one state give the launch more
// blocks (12 qubits complex128, 12 observables: 0.63 ms per Jacobian against
// 0.74 with chunks of 3). PLK_RTILE_NJ overrides, up to RT_NJMAX.
int rtile_chunk(int dtype, int n, int num_cus) {
    const int e = env_int("PLK_RTILE_NJ", 0);
    if (e > 0) {
        return std::min(RT_NJMAX, e);
    }
    const int T = (dtype == PLK_C64 ? 4 : 3) + RT_LANE_BITS;
    const bool few_tiles = n - T < 30 && (int64_t{1} << std::max(0, n - T)) < num_cus;
    return few_tiles ? 1 : 3;
}

// Register-tile passes serve every state of at least one register tile:
// with their observable chunks side by side they measured faster than the
// LDS tile passes at every size from 10 to 21 qubits in both precisions
// (6-layer adjoint, 9 observables, complex128: 0.70 vs 2.52 ms at 10 qubits,
// 1.32 vs 2.12 at 16, 8.35 vs 11.26 at 21). PLK_RTILE=0 turns them off.
bool use_rtile(int n, int dtype) {
    const int T = rt_register_bits(dtype) + RT_LANE_BITS;
    return env_int("PLK_RTILE", -1) != 0 && n >= T;
}

void apply_single(Device &dev, const State &s, const CompiledOp &op);

// Launches a forward plan whose tables are at pl (no host synchronisation
// unless a step is a single op outside the fused kernels).
void enqueue_forward(Device &dev, const State &s, const std::vector<Step> &steps,
                     const PlanImage &img, const PlanPlace &pl) {
    for (size_t i = 0; i < steps.size(); i++) {
        if (steps[i].tiled && steps[i].pass.rt) {
            RtLaunch L = img.progs[i].L;
            L.nj = 0;
            L.write_phi = 1;
            L.nslots = 0;
            launch_rtile(dev, s, nullptr, nullptr, pl.rt + img.progs[i].off, L);
        } else if (steps[i].fused) {
            SegLaunch L = img.metas[i].L;
            L.write_phi = 1;
            L.nj = 0;
            L.nslots = 0;
            launch_segment(dev, s, nullptr, pl.ops + img.metas[i].first, L);
        } else if (steps[i].tiled) {
            TileLaunch TL = img.metas[i].TL;
            TL.nj = 0;
            TL.write_phi = 1;
            TL.nslots = 0;
            launch_tile(dev, s, nullptr, pl.bundles + img.metas[i].first_bundle, pl.ops, TL);
        } else if (img.singles[i].ok) {
            StepLaunch L = img.singles[i].L;
            L.nj = 0;
            L.trainable = 0;
            L.write_phi = 1;
            launch_step(dev, s, nullptr, pl.mats + img.singles[i].off, L);
        } else {
            apply_single(dev, s, steps[i].single.U);
        }
    }
}

void run_forward(Device &dev, const State &s, const std::vector<Step> &steps) {
    const PlanImage img = build_image(steps, s);
    enqueue_forward(dev, s, steps, img, place_in_arena(dev, img));
    // The host image is read by the stream's copies.
    PLK_HIP(hipStreamSynchronize(dev.stream));
}

void apply_single(Device &dev, const State &s, const CompiledOp &op) {
    switch (op.kind) {
    case OpKind::Identity:
        return;
    case OpKind::Parity:
        launch_parity(dev, s, op.wires, op.even, op.odd);
        return;
    case OpKind::Dense:
        if (op.wires.size() <= 2) {
            PlanOp p;
            p.U = op;
            std::vector<PlanOp> v;
            v.push_back(std::move(p));
            run_forward(dev, s, plan_steps(std::move(v)));
        } else {
            launch_dense(dev, s, op);
        }
        return;
    }
}

void apply_generator_dev(Device &dev, const State &s, const CompiledGen &g) {
    if (g.kind == OpKind::Parity) {
        launch_parity(dev, s, g.wires, 1.0, -1.0);
    } else {
        launch_dense_matrix(dev, s, g.wires, g.mat);
    }
}

// ---------------------------------------------------------------------------
// Observable application and expectation values.
// ---------------------------------------------------------------------------
void apply_obs_general(Device &dev, const State &s, const ObsNode &o);

const PauliDevice *ObsNode::pauli_on(Device &dev, int n) const {
    auto it = pauli_dev.find({dev.id, n});
    if (it != pauli_dev.end()) {
        return it->second.get();
    }
    std::vector<PauliTerm> terms;
    if (!to_pauli(*this, n, terms)) {
        return nullptr;
    }
    auto &slot = pauli_dev[{dev.id, n}];
    slot = pauli_upload(dev, terms);
    return slot.get();
}

void check_sparse_wires(const ObsNode &o, const State &s) {
    if (o.wires.size() != static_cast<size_t>(s.n)) {
        raise(PLK_ERR_INVALID, "SparseH wire count does not match state-vector size");
    }
}

void apply_obs_into(Device &dev, const State &out, const State &in, const ObsNode &o) {
    if (const PauliDevice *t = o.pauli_on(dev, in.n)) {
        launch_pauli_apply(dev, out, in, *t);
        return;
    }
    if (o.kind == ObsNode::Sparse) {
        check_sparse_wires(o, in);
        csr_spmv(dev, out, in, o.csr_on(dev, in));
        return;
    }
    launch_copy(dev, out, in);
    apply_obs_general(dev, out, o);
}

void apply_obs_general(Device &dev, const State &s, const ObsNode &o) {
    switch (o.kind) {
    case ObsNode::Named:
        apply_single(dev, s, compile_op(o.name, o.wires, false, {}, {}, s.n));
        return;
    case ObsNode::Hermitian: {
        CompiledOp op;
        op.kind = OpKind::Dense;
        op.wires = o.wires;
        op.mat = o.matrix;
        apply_single(dev, s, op);
        return;
    }
    case ObsNode::Tensor:
        for (auto &t : o.terms) {
            apply_obs_general(dev, s, *t);
        }
        return;
    case ObsNode::Hamiltonian: {
        TmpState buf(s.n, s.dtype, s.device);
        TmpState tmp(s.n, s.dtype, s.device);
        launch_fill_zero(dev, buf.s);
        for (size_t t = 0; t < o.terms.size(); t++) {
            apply_obs_into(dev, tmp.s, s, *o.terms[t]);
            launch_axpy(dev, buf.s, o.coeffs[t], tmp.s);
        }
        launch_copy(dev, s, buf.s);
        PLK_HIP(hipStreamSynchronize(dev.stream));
        return;
    }
    case ObsNode::Sparse: {
        check_sparse_wires(o, s);
        TmpState tmp(s.n, s.dtype, s.device);
        csr_spmv(dev, tmp.s, s, o.csr_on(dev, s));
        launch_copy(dev, s, tmp.s);
        PLK_HIP(hipStreamSynchronize(dev.stream));
        return;
    }
    }
}

double expval_obs(Device &dev, const State &s, const ObsNode &o) {
    if (const PauliDevice *t = o.pauli_on(dev, s.n)) {
        return pauli_expval(dev, s, *t);
    }
    if (o.kind == ObsNode::Sparse) {
        // sum_r Re(conj(psi_r) (A psi)_r) without materialising A psi.
        check_sparse_wires(o, s);
        return csr_expval(dev, s, o.csr_on(dev, s));
    }
    if (o.kind == ObsNode::Hermitian) {
        // <psi|M|psi> group by group, without a copy of the state.
        return dense_expval(dev, s, o.wires, o.matrix);
    }
    TmpState tmp(s.n, s.dtype, s.device);
    launch_copy(dev, tmp.s, s);
    apply_obs_general(dev, tmp.s, o);
    const State *xs[1] = {&s};
    double ri[2];
    inner_products(dev, xs, 1, tmp.s, ri);
    return ri[0];
}

// ---------------------------------------------------------------------------
// Operation list (AdjointDiffKokkos.hpp:17-173).
// ---------------------------------------------------------------------------
struct OpsData {
    std::vector<std::string> names;
    std::vector<std::vector<double>> params;
    std::vector<std::vector<size_t>> wires;
    std::vector<bool> inverses;
    std::vector<std::vector<cd>> matrices;
    size_t num_par_ops() const {
        size_t c = 0;
        for (auto &p : params) {
            c += p.empty() ? 0 : 1;
        }
        return c;
    }
};

// Forward plan of the whole op list (applyOperations, AdjointDiffKokkos.hpp:213-226).
std::vector<Step> forward_plan(const OpsData &ops, int n, int dtype) {
    std::vector<PlanOp> v;
    for (size_t i = 0; i < ops.names.size(); i++) {
        PlanOp p;
        p.U = compile_op(ops.names[i], ops.wires[i], ops.inverses[i], ops.params[i],
                         ops.matrices[i], n);
        v.push_back(std::move(p));
    }
    if (use_rtile(n, dtype)) {
        return plan_tiles(std::move(v), n, rt_register_bits(dtype) + RT_LANE_BITS, low_bits(dtype),
                          RT_PMAX, false, rt_register_bits(dtype) + 2);
    }
    const int T = tile_bits_forward(n, dtype);
    if (T > 0) {
        return plan_tiles(std::move(v), n, T, low_bits(dtype));
    }
    return plan_steps(std::move(v));
}

// The reverse walk of AdjointDiffKokkos.hpp:432-470: U^dagger of every op
// after the last trainable one, in reverse order, each trainable op carrying
// its generator, Jacobian column and coefficient -2 s (inverse ? -1 : 1).
// StatePrep / BasisState are skipped; ops with several parameters abort.
std::vector<PlanOp> reverse_ops(const OpsData &ops, const std::vector<size_t> &tp, int n) {
    const size_t n_tp = tp.size();
    std::vector<PlanOp> rev;
    size_t trainable_number = n_tp - 1;
    size_t current_param_idx = ops.num_par_ops() - 1;
    auto tp_it = tp.rbegin();
    for (long op_idx = static_cast<long>(ops.names.size()) - 1; op_idx >= 0; op_idx--) {
        const size_t o = static_cast<size_t>(op_idx);
        if (ops.params[o].size() > 1) {
            raise(PLK_ERR_INVALID,
                  "The operation is not supported using the adjoint "
                  "differentiation method");
        }
        if (ops.names[o] == "StatePrep" || ops.names[o] == "BasisState") {
            continue;
        }
        if (tp_it == tp.rend()) {
            break;
        }
        PlanOp p;
        p.U = compile_op(ops.names[o], ops.wires[o], !ops.inverses[o], ops.params[o],
                         ops.matrices[o], n);
        p.src = static_cast<int32_t>(o);
        p.adj = !ops.inverses[o];
        if (!ops.params[o].empty()) {
            if (current_param_idx == *tp_it) {
                p.trainable = true;
                p.G = compile_generator(ops.names[o], ops.wires[o], n);
                p.coef = -2.0 * p.G.scale * (ops.inverses[o] ? -1.0 : 1.0);
                p.col = static_cast<int>(trainable_number);
                trainable_number--;
                ++tp_it;
            }
            current_param_idx--;
        }
        rev.push_back(std::move(p));
    }
    return rev;
}

// 64-bit structure hash (multiply-xorshift over 64-bit words).
struct Hasher {
    uint64_t h = 0x243F6A8885A308D3ull;
    void add(uint64_t v) {
        h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        h *= 0xFF51AFD7ED558CCDull;
        h ^= h >> 32;
    }
    void add(double v) {
        uint64_t b;
        std::memcpy(&b, &v, sizeof b);
        add(b);
    }
    void add(const std::string &str) {
        add(static_cast<uint64_t>(str.size()));
        add(static_cast<uint64_t>(std::hash<std::string>{}(str)));
    }
};

// ---------------------------------------------------------------------------
// Adjoint Jacobian (AdjointDiffKokkos.hpp:404-478).
//
// The reverse walk over the operations, the trainable-parameter bookkeeping
// and the early stop follow the reference line by line; the per-operation
// work (mu copy, U^dagger on lambda, generator on mu, n_obs inner products,
// U^dagger on every H lambda) is replaced by fused passes.
// ---------------------------------------------------------------------------

// Launches reverse step i (a fused segment, register-tile or LDS-tile pass
// over lambda and every H_j lambda, with its Jacobian reductions); returns
// false for a single op, which the caller runs.
// The adjoint's lambda. Register-tile passes run their observable chunks
// side by side, all reading cur, and write the updated lambda to alt; the
// two then swap. Without alt (one chunk) they update cur in place.
struct PhiPair {
    State cur, alt;
};

bool enqueue_reverse_step(Device &dev, PhiPair &phi, void *const *hptr, size_t n_obs,
                          size_t n_tp, const Step &st, const PlanImage &img, size_t i,
                          const PlanPlace &pl, double *d_jac) {
    const State &lambda = phi.cur;
    if (st.fused) {
        const SegMeta &m = img.metas[i];
        const int njmax = std::min(SEG_NJMAX, seg_max_nj(lambda.dtype, m.L.K));
        size_t j0 = 0;
        while (j0 < n_obs) {
            const size_t left = n_obs - j0;
            int nj = 1;
            for (int c : {8, 4, 2, 1}) {
                if (c <= njmax && static_cast<size_t>(c) <= left) {
                    nj = c;
                    break;
                }
            }
            SegLaunch L = m.L;
            L.nj = nj;
            L.write_phi = (j0 + nj == n_obs) ? 1 : 0;
            const int grid = launch_segment(dev, lambda, hptr + j0, pl.ops + m.first, L);
            if (L.nslots > 0) {
                launch_segment_reduce(dev, grid, L.nslots, nj, static_cast<int>(j0), m.cols,
                                      m.coefs, d_jac, static_cast<int>(n_tp));
            }
            j0 += nj;
        }
        return true;
    }
    if (st.tiled && st.pass.rt) {
        // The chunks (balanced: widths nj and nj - 1, wider first) run side
        // by side in one launch, up to RT_OBS_MAX states, when lambda has a
        // second buffer.
        const std::vector<int> sizes = chunk_sizes(n_obs, rtile_chunk(lambda.dtype, lambda.n, dev.num_cus));
        size_t j0 = 0, ci = 0;
        while (ci < sizes.size()) {
            const int nj = sizes[ci];
            int nc = 1, nbig = 1;
            size_t width = static_cast<size_t>(nj);
            while (phi.alt.d != nullptr && ci + nc < sizes.size() &&
                   width + static_cast<size_t>(sizes[ci + nc]) <= static_cast<size_t>(RT_OBS_MAX)) {
                width += static_cast<size_t>(sizes[ci + nc]);
                nbig += sizes[ci + nc] == nj ? 1 : 0;
                nc++;
            }
            RtLaunch L = img.progs[i].L;
            L.nj = nj;
            L.nc = nc;
            L.nbig = nbig;
            L.write_phi = (j0 + width == n_obs) ? 1 : 0;
            const bool swap = L.write_phi && phi.alt.d != nullptr;
            const int grid = launch_rtile(dev, phi.cur, swap ? &phi.alt : nullptr, hptr + j0,
                                          pl.rt + img.progs[i].off, L);
            if (L.nslots > 0) {
                launch_tile_reduce(dev, grid, L.nslots, static_cast<int>(width), static_cast<int>(j0),
                                   st.pass.cols.data(), st.pass.coefs.data(), d_jac,
                                   static_cast<int>(n_tp));
            }
            if (swap) {
                std::swap(phi.cur, phi.alt);
            }
            j0 += width;
            ci += nc;
        }
        return true;
    }
    if (!st.fused && !st.tiled && img.singles[i].ok) {
        const StepMeta &m = img.singles[i];
        size_t j0 = 0;
        for (int nj : chunk_sizes(n_obs, SEG_NJMAX)) {
            StepLaunch L = m.L;
            L.nj = nj;
            L.write_phi = (j0 + nj == n_obs) ? 1 : 0;
            const int grid = launch_step(dev, lambda, hptr + j0, pl.mats + m.off, L);
            if (L.trainable) {
                launch_tile_reduce(dev, grid, 1, nj, static_cast<int>(j0), &m.col, &m.coef, d_jac,
                                   static_cast<int>(n_tp));
            }
            j0 += nj;
        }
        return true;
    }
    if (st.tiled) {
        const SegMeta &m = img.metas[i];
        size_t j0 = 0;
        for (int nj : chunk_sizes(n_obs, tile_max_nj(lambda.dtype, m.TL.T))) {
            TileLaunch TL = m.TL;
            TL.nj = nj;
            TL.write_phi = (j0 + nj == n_obs) ? 1 : 0;
            const int grid = launch_tile(dev, lambda, hptr + j0, pl.bundles + m.first_bundle,
                                         pl.ops, TL);
            if (TL.nslots > 0) {
                launch_tile_reduce(dev, grid, TL.nslots, nj, static_cast<int>(j0),
                                   st.pass.cols.data(), st.pass.coefs.data(), d_jac,
                                   static_cast<int>(n_tp));
            }
            j0 += nj;
        }
        return true;
    }
    return false;
}

// A captured sweep: copy of the input state, forward passes, observables,
// Jacobian zero fill and reverse passes of one circuit structure, replayed
// with a fresh upload of the plan and Pauli tables (angles and coefficients
// live there). The graph owns every buffer its kernels touch except the
// input state and the device's partials buffer (checked before each replay).
struct SweepGraph {
    uint64_t key = 0;
    int device = -1;
    hipStream_t stream = nullptr;
    const void *ref = nullptr;
    std::vector<unsigned char> sig;
    std::unique_ptr<TmpState> lambda, lambda2; // lambda2: second buffer of PhiPair (or none)
    std::vector<std::unique_ptr<TmpState>> H;
    std::vector<void *> hptr;
    void *tables = nullptr;
    size_t fwd_bytes = 0, rev_bytes = 0;
    double *jac = nullptr;
    double *red = nullptr;
    hipGraphExec_t exec = nullptr;
    ~SweepGraph() {
        (void)hipSetDevice(device);
        (void)hipStreamSynchronize(stream);
        if (exec) {
            (void)hipGraphExecDestroy(exec);
        }
        if (tables) {
            (void)hipFree(tables);
        }
        if (jac) {
            (void)hipFree(jac);
        }
    }
};

constexpr int kGraphMaxQubits = 18;
constexpr size_t kGraphsPerDevice = 128;
constexpr size_t kGraphBytesPerDevice = size_t{4} << 30; // states held by graphs

std::map<int, std::vector<std::unique_ptr<SweepGraph>>> &sweep_graphs() {
    static std::map<int, std::vector<std::unique_ptr<SweepGraph>>> g; // released with the process
    return g;
}

// Structures seen once (a capture pays off from the second call on).
std::map<int, std::map<uint64_t, int>> &sweep_seen() {
    static std::map<int, std::map<uint64_t, int>> seen;
    return seen;
}

size_t graph_bytes(const SweepGraph &g) {
    return (1 + g.H.size() + (g.lambda2 ? 1 : 0)) * g.lambda->s.bytes();
}

// Whether the reverse plan runs register-tile chunks side by side, which
// needs lambda in two buffers.
bool needs_alt(const Device &dev, const std::vector<Step> &steps, size_t n_obs, int dtype, int n) {
    if (n_obs <= static_cast<size_t>(rtile_chunk(dtype, n, dev.num_cus))) {
        return false;
    }
    for (const auto &st : steps) {
        if (st.tiled && st.pass.rt) {
            return true;
        }
    }
    return false;
}

// Everything but the angles: op names / wires / inverse flags / parameter
// counts / matrix sizes, the trainable set, sizes and flags.
uint64_t sweep_key(const State &ref, const OpsData &ops, const std::vector<size_t> &tp,
                   size_t n_obs, bool apply_ops) {
    Hasher H;
    H.add(static_cast<uint64_t>(ref.n));
    H.add(static_cast<uint64_t>(ref.dtype));
    H.add(static_cast<uint64_t>(n_obs));
    H.add(static_cast<uint64_t>(apply_ops ? 1 : 0));
    H.add(static_cast<uint64_t>(ops.names.size()));
    for (size_t i = 0; i < ops.names.size(); i++) {
        H.add(ops.names[i]);
        H.add(static_cast<uint64_t>(ops.wires[i].size() | (ops.inverses[i] ? size_t{1} << 40 : 0)));
        for (auto w : ops.wires[i]) {
            H.add(static_cast<uint64_t>(w));
        }
        H.add(static_cast<uint64_t>(ops.params[i].size()));
        H.add(static_cast<uint64_t>(ops.matrices[i].size()));
    }
    H.add(static_cast<uint64_t>(tp.size()));
    for (auto t : tp) {
        H.add(static_cast<uint64_t>(t));
    }
    return H.h;
}

// Stream capture with profiling scopes off; always ends the capture.
struct CaptureGuard {
    hipStream_t stream;
    bool open = true;
    explicit CaptureGuard(hipStream_t s) : stream(s) {
        set_capturing(true);
        PLK_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
    }
    hipGraph_t end() {
        hipGraph_t g = nullptr;
        open = false;
        set_capturing(false);
        PLK_HIP(hipStreamEndCapture(stream, &g));
        return g;
    }
    ~CaptureGuard() {
        if (open) {
            hipGraph_t g = nullptr;
            set_capturing(false);
            if (hipStreamEndCapture(stream, &g) == hipSuccess && g) {
                (void)hipGraphDestroy(g);
            }
        }
    }
};

// Whether a plan holds a single-op step outside the step kernels (those
// synchronise with the host, so the plan cannot be captured).
bool steps_have_single(const std::vector<Step> &steps, int n) {
    for (const auto &st : steps) {
        if (!st.fused && !st.tiled && !step_kernel_ok(st.single, n)) {
            return true;
        }
    }
    return false;
}

void adjoint_jacobian(const State &ref, const std::vector<std::shared_ptr<const ObsNode>> &obs,
                      const OpsData &ops, const std::vector<size_t> &tp, bool apply_ops,
                      double *jac) {
    if (tp.empty()) {
        raise(PLK_ERR_INVALID, "No trainable parameters provided.");
    }
    const size_t n_obs = obs.size();
    const size_t n_tp = tp.size();
    std::fill(jac, jac + n_obs * n_tp, 0.0);
    Device &dev = dev_of(ref);
    const int n = ref.n;

    // Reverse plan first: it validates the op list before any device work.
    std::unique_ptr<HostScope> plan_scope = std::make_unique<HostScope>("host_adjoint_plan");
    std::vector<PlanOp> rev = reverse_ops(ops, tp, n);
    if (n_obs == 0) {
        return;
    }
    const bool rt = use_rtile(n, ref.dtype);
    const int T_adj = rt ? rt_register_bits(ref.dtype) + RT_LANE_BITS : tile_bits_adjoint(n, ref.dtype);
    const std::vector<Step> steps =
        rt ? plan_tiles(std::move(rev), n, T_adj, low_bits(ref.dtype), RT_PMAX, false,
                        rt_register_bits(ref.dtype) + 2)
           : T_adj > 0 ? plan_tiles(std::move(rev), n, T_adj, low_bits(ref.dtype))
                       : plan_steps(std::move(rev));
    const std::vector<Step> fwd = apply_ops ? forward_plan(ops, n, ref.dtype) : std::vector<Step>{};
    const PlanImage img_f = build_image(fwd, ref);
    const PlanImage img_r = build_image(steps, ref);
    plan_scope.reset();

    // Small circuits run as a replayed HIP graph: their sweeps are hundreds
    // of launches of a few microseconds each.
    // Observables enter the graph as Pauli tables copied with the plan
    // tables, so circuits that differ only in angles or in observable
    // coefficients share one graph.
    struct PauliImg {
        std::vector<unsigned char> bytes;
        size_t toff = 0;
        int ngroups = 0;
    };
    std::vector<PauliImg> pimg;
    bool graph = n <= kGraphMaxQubits && env_int("PLK_GRAPH", 1) != 0 &&
                 !steps_have_single(steps, n) && !steps_have_single(fwd, n);
    for (size_t j = 0; graph && j < n_obs; j++) {
        std::vector<PauliTerm> terms;
        graph = to_pauli(*obs[j], n, terms);
        PauliImg pi;
        if (graph) {
            pi.bytes = pauli_image(terms, pi.toff, pi.ngroups);
        }
        pimg.push_back(std::move(pi));
    }
    const uint64_t key = graph ? sweep_key(ref, ops, tp, n_obs, apply_ops) : 0;
    if (graph) {
        auto &seen = sweep_seen()[dev.id];
        if (seen.size() > 65536) {
            seen.clear();
        }
        graph = ++seen[key] >= 2;
    }
    if (graph) {
        std::unique_ptr<HostScope> phase = std::make_unique<HostScope>("host_adjoint_graph");
        std::vector<unsigned char> sig;
        append_signature(img_f, sig);
        append_signature(img_r, sig);
        size_t pauli_bytes = 0;
        for (const auto &pi : pimg) {
            const uint64_t f[3] = {pi.bytes.size(), pi.toff, static_cast<uint64_t>(pi.ngroups)};
            const auto *b = reinterpret_cast<const unsigned char *>(f);
            sig.insert(sig.end(), b, b + sizeof f);
            pauli_bytes += (pi.bytes.size() + 255) & ~size_t{255};
        }
        auto &list = sweep_graphs()[dev.id];
        SweepGraph *g = nullptr;
        for (auto it = list.begin(); it != list.end(); ++it) {
            SweepGraph &c = **it;
            const bool same = c.key == key && c.ref == ref.d;
            if (same) {
                if (c.sig != sig || c.red != dev.red) {
                    list.erase(it); // stale layout or partials buffer: capture again
                } else {
                    std::rotate(it, it + 1, list.end()); // most recently used last
                    g = list.back().get();
                }
                break;
            }
        }
        if (g == nullptr) {
            // Least recently used first in the list: evict from the front.
            size_t held = 0;
            for (const auto &c : list) {
                held += graph_bytes(*c);
            }
            const size_t need = (2 + n_obs) * ref.bytes();
            while (!list.empty() &&
                   (list.size() >= kGraphsPerDevice || held + need > kGraphBytesPerDevice)) {
                held -= graph_bytes(*list.front());
                list.erase(list.begin());
            }
            auto ng = std::make_unique<SweepGraph>();
            ng->key = key;
            ng->device = dev.id;
            ng->stream = dev.stream;
            ng->ref = ref.d;
            ng->sig = sig;
            ng->lambda = std::make_unique<TmpState>(n, ref.dtype, ref.device);
            if (needs_alt(dev, steps, n_obs, ref.dtype, n)) {
                ng->lambda2 = std::make_unique<TmpState>(n, ref.dtype, ref.device);
            }
            for (size_t j = 0; j < n_obs; j++) {
                ng->H.push_back(std::make_unique<TmpState>(n, ref.dtype, ref.device));
                ng->hptr.push_back(ng->H.back()->s.d);
            }
            ng->fwd_bytes = image_bytes(img_f);
            ng->rev_bytes = image_bytes(img_r);
            ng->tables = dev_malloc(dev.id, ng->fwd_bytes + ng->rev_bytes + pauli_bytes + 256,
                                    "graph tables");
            ng->jac = static_cast<double *>(dev_malloc(dev.id, sizeof(double) * n_obs * n_tp, "graph Jacobian"));
            // The launchers grow the partials buffer on demand, which would
            // synchronise inside the capture: size it for any pass first.
            ensure_red(dev,
                       static_cast<size_t>(dev.num_cus) * 32 * std::max(TILE_PMAX, RT_PMAX) *
                           std::max(RT_NJMAX, SEG_NJMAX),
                       256);
            auto *tb = static_cast<unsigned char *>(ng->tables);
            // Addresses only: the tables are copied before every replay.
            const PlanPlace pf = place_addresses(img_f, tb);
            const PlanPlace pr = place_addresses(img_r, tb + ng->fwd_bytes);
            {
                CaptureGuard cap(dev.stream);
                const State &lam = ng->lambda->s;
                launch_copy(dev, lam, ref);
                enqueue_forward(dev, lam, fwd, img_f, pf);
                unsigned char *pt = tb + ng->fwd_bytes + ng->rev_bytes;
                for (size_t j = 0; j < n_obs; j++) {
                    launch_pauli_apply_at(dev, ng->H[j]->s, lam, pt, pimg[j].toff, pimg[j].ngroups);
                    pt += (pimg[j].bytes.size() + 255) & ~size_t{255};
                }
                launch_fill_zero_doubles(dev, ng->jac, n_obs * n_tp);
                PhiPair phi{lam, ng->lambda2 ? ng->lambda2->s : State{}};
                for (size_t i = 0; i < steps.size(); i++) {
                    enqueue_reverse_step(dev, phi, ng->hptr.data(), n_obs, n_tp, steps[i], img_r,
                                         i, pr, ng->jac);
                }
                hipGraph_t graph_def = cap.end();
                const hipError_t e = hipGraphInstantiate(&ng->exec, graph_def, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph_def);
                PLK_HIP(e);
            }
            ng->red = dev.red;
            list.push_back(std::move(ng));
            g = list.back().get();
        }
        auto *tb = static_cast<unsigned char *>(g->tables);
        copy_image(dev, img_f, tb);
        copy_image(dev, img_r, tb + g->fwd_bytes);
        unsigned char *pt = tb + g->fwd_bytes + g->rev_bytes;
        for (const auto &pi : pimg) {
            if (!pi.bytes.empty()) {
                PLK_HIP(hipMemcpyAsync(pt, pi.bytes.data(), pi.bytes.size(), hipMemcpyHostToDevice,
                                       dev.stream));
            }
            pt += (pi.bytes.size() + 255) & ~size_t{255};
        }
        PLK_HIP(hipGraphLaunch(g->exec, dev.stream));
        PLK_HIP(hipMemcpyAsync(jac, g->jac, sizeof(double) * n_obs * n_tp, hipMemcpyDeviceToHost,
                               dev.stream));
        // Also keeps the host images alive until their copies have run.
        PLK_HIP(hipStreamSynchronize(dev.stream));
        return;
    }

    TmpState lambda(n, ref.dtype, ref.device);
    launch_copy(dev, lambda.s, ref);
    if (apply_ops) {
        enqueue_forward(dev, lambda.s, fwd, img_f, place_in_arena(dev, img_f));
    }
    std::vector<std::unique_ptr<TmpState>> H;
    for (size_t j = 0; j < n_obs; j++) {
        H.push_back(std::make_unique<TmpState>(n, ref.dtype, ref.device));
        apply_obs_into(dev, H.back()->s, lambda.s, *obs[j]);
    }
    std::vector<void *> hptr(n_obs);
    std::vector<const State *> hstate(n_obs);
    for (size_t j = 0; j < n_obs; j++) {
        hptr[j] = H[j]->s.d;
        hstate[j] = &H[j]->s;
    }

    double *d_jac = ensure_jac(dev, n_obs * n_tp);
    launch_fill_zero_doubles(dev, d_jac, n_obs * n_tp);
    std::vector<std::pair<size_t, double>> host_entries; // (flat index, value)

    const PlanPlace pl = place_in_arena(dev, img_r);
    std::unique_ptr<TmpState> mu, lambda2;
    if (needs_alt(dev, steps, n_obs, ref.dtype, n)) {
        lambda2 = std::make_unique<TmpState>(n, ref.dtype, ref.device);
    }
    PhiPair phi{lambda.s, lambda2 ? lambda2->s : State{}};
    for (size_t i = 0; i < steps.size(); i++) {
        const Step &st = steps[i];
        if (enqueue_reverse_step(dev, phi, hptr.data(), n_obs, n_tp, st, img_r, i, pl, d_jac)) {
            continue;
        }
        const PlanOp &p = st.single;
        if (p.trainable) {
            if (!mu) {
                mu = std::make_unique<TmpState>(n, ref.dtype, ref.device);
            }
            launch_copy(dev, mu->s, phi.cur);
        }
        apply_single(dev, phi.cur, p.U);
        if (p.trainable) {
            apply_generator_dev(dev, mu->s, p.G);
            std::vector<double> ri(2 * n_obs);
            inner_products(dev, hstate.data(), static_cast<int>(n_obs), mu->s, ri.data());
            for (size_t j = 0; j < n_obs; j++) {
                host_entries.push_back({j * n_tp + static_cast<size_t>(p.col),
                                        p.coef * ri[2 * j + 1]});
            }
        }
        for (size_t j = 0; j < n_obs; j++) {
            apply_single(dev, H[j]->s, p.U);
        }
    }
    PLK_HIP(hipMemcpyAsync(jac, d_jac, sizeof(double) * n_obs * n_tp, hipMemcpyDeviceToHost,
                           dev.stream));
    // Also keeps the host images alive until their copies have run.
    PLK_HIP(hipStreamSynchronize(dev.stream));
    for (auto &e : host_entries) {
        jac[e.first] = e.second;
    }
}

// ---------------------------------------------------------------------------
// Batched circuits: independent circuits, each from |0...0> with one
// observable, planned on the host into LDS tile passes and run by one
// persistent launch (kernels_batch.hip).
// ---------------------------------------------------------------------------
struct BatchInput {
    int n = 0;
    const OpsData *ops = nullptr;
    const ObsNode *obs = nullptr;
    std::vector<size_t> tp;
};

int batch_tile_bits(int dtype) {
    const int T = env_int("PLK_BATCH_TILE_BITS", dtype == PLK_C64 ? 12 : 11);
    return std::max(low_bits(dtype) + 2, std::min(T, BATCH_TMAX));
}

// Host-side batch plan; every table is uploaded as one blob.
template <typename T> struct BatchPlan {
    std::vector<BatchCircuit> circuits;
    std::vector<BatchPass> passes;
    std::vector<TileBundle> bundles;
    std::vector<unsigned char> ops; // packed op groups
    std::vector<BatchTerm> terms;
    std::vector<int32_t> slot_cols;
    std::vector<double> slot_coefs;
    std::vector<double> cost;
    std::vector<BatchPatch> patches; // offsets into ops
    int64_t jac_len = 0;
    int64_t nparams = 0; // parametric ops of the planned circuits

    // Releases dst's memory (a merged part is no longer needed).
    void swap_into(BatchPlan &dst) {
        std::swap(*this, dst);
    }
};

// Patched gate of an op: one-parameter rotations without an explicit matrix.
int16_t patch_gate_of(const OpsData &ops, size_t i) {
    if (ops.params[i].size() != 1 || !ops.matrices[i].empty()) {
        return BPG_NONE;
    }
    const std::string &nm = ops.names[i];
    if (nm == "RX") {
        return BPG_RX;
    }
    if (nm == "RY") {
        return BPG_RY;
    }
    if (nm == "RZ") {
        return BPG_RZ;
    }
    if (nm == "PhaseShift") {
        return BPG_PHASE;
    }
    return BPG_NONE;
}

template <typename T>
void plan_batch_passes(BatchPlan<T> &bp, std::vector<Step> &&steps, const BatchCircuit &c,
                       int Tt, bool forward, int &first, int &count) {
    first = static_cast<int>(bp.passes.size());
    if (forward && steps.empty() && !c.resident) {
        Step s;
        s.tiled = true;
        for (int b = 0; b < Tt; b++) {
            s.pass.tb.push_back(b);
        }
        steps.push_back(std::move(s));
    }
    for (auto &st : steps) {
        if (!st.tiled) {
            raise(PLK_ERR_UNSUPPORTED,
                  "batched circuits support dense gates and generators on at most two wires");
        }
        SegMeta meta;
        BatchPass P{};
        P.bundle_first = static_cast<int32_t>(bp.bundles.size());
        build_pass<T>(st.pass, c.n, bp.ops, bp.bundles, meta, &bp.patches);
        P.T = meta.TL.T;
        P.bundle_count = meta.TL.nbundles;
        P.nslots = st.pass.nslots;
        P.slot_first = static_cast<int32_t>(bp.slot_cols.size());
        for (int q = 0; q < P.T; q++) {
            P.tb[q] = st.pass.tb[q];
            P.tmask |= 1u << st.pass.tb[q];
        }
        for (int k = 0; k < P.nslots; k++) {
            bp.slot_cols.push_back(st.pass.cols[k]);
            bp.slot_coefs.push_back(st.pass.coefs[k]);
        }
        size_t ops_in_pass = 0;
        for (const auto &b : st.pass.bundles) {
            ops_in_pass += b.ops.size();
        }
        const double amps = std::ldexp(1.0, c.n);
        bp.cost.back() += amps * (static_cast<double>(P.bundle_count) + 0.25 * ops_in_pass) *
                              (forward ? 1.0 : 2.0) +
                          (c.resident ? 0.0 : 8.0 * amps);
        bp.passes.push_back(P);
    }
    count = static_cast<int>(bp.passes.size()) - first;
    if (!c.resident && count > 0) {
        if (forward) {
            bp.passes[first].flags |= BP_INIT;
        } else {
            bp.passes.back().flags |= BP_NO_STORE;
        }
    }
}

template <typename T>
void plan_batch_circuit(BatchPlan<T> &bp, const BatchInput &in, int tile_bits, int L) {
    const int n = in.n;
    BatchCircuit c{};
    c.n = n;
    c.resident = n <= tile_bits ? 1 : 0;
    const int Tt = c.resident ? n : tile_bits;
    const int Lc = c.resident ? 0 : L;
    c.jac_offset = bp.jac_len;
    bp.cost.push_back(0.0);

    std::vector<PauliTerm> pt;
    if (!to_pauli(*in.obs, n, pt)) {
        raise(PLK_ERR_UNSUPPORTED,
              "batched circuits support observables that are sums of Pauli words");
    }
    c.term_first = static_cast<int32_t>(bp.terms.size());
    c.term_count = static_cast<int32_t>(pt.size());
    for (const auto &t : pt) {
        bp.terms.push_back(BatchTerm{t.x, t.zy, t.coeff.real(), t.coeff.imag()});
    }
    bp.cost.back() += std::ldexp(1.0, n) * (1.0 + 0.25 * static_cast<double>(pt.size()));

    const OpsData &ops = *in.ops;
    std::vector<PlanOp> rev = reverse_ops(ops, in.tp, n);
    std::vector<PlanOp> fwd;
    for (size_t i = 0; i < ops.names.size(); i++) {
        PlanOp p;
        p.U = compile_op(ops.names[i], ops.wires[i], ops.inverses[i], ops.params[i],
                         ops.matrices[i], n);
        fwd.push_back(std::move(p));
    }
    // Global parameter index of every parametric op (circuit order), and the
    // patch gate of the ones whose U payload is rewritten per call.
    std::vector<int32_t> pidx(ops.names.size(), -1);
    int64_t np = bp.nparams;
    for (size_t i = 0; i < ops.names.size(); i++) {
        if (!ops.params[i].empty()) {
            pidx[i] = static_cast<int32_t>(np++);
        }
    }
    if (np > INT32_MAX) {
        raise(PLK_ERR_UNSUPPORTED, "batched parameters exceed 2^31");
    }
    auto mark = [&](PlanOp &p) {
        if (p.src >= 0 && pidx[p.src] >= 0) {
            p.param = pidx[p.src];
            p.patch_gate = patch_gate_of(ops, static_cast<size_t>(p.src));
        }
    };
    for (size_t i = 0; i < fwd.size(); i++) {
        fwd[i].src = static_cast<int32_t>(i);
        fwd[i].adj = ops.inverses[i];
        mark(fwd[i]);
    }
    for (auto &p : rev) {
        mark(p);
    }
    bp.nparams = np;
    plan_batch_passes(bp, plan_tiles(std::move(fwd), n, Tt, Lc, BATCH_PMAX), c, Tt, true,
                      c.fwd_first, c.fwd_count);
    plan_batch_passes(bp, plan_tiles(std::move(rev), n, Tt, Lc, BATCH_PMAX), c, Tt, false,
                      c.bwd_first, c.bwd_count);
    bp.jac_len += static_cast<int64_t>(in.tp.size());
    bp.circuits.push_back(c);
}

struct Blob {
    std::vector<unsigned char> bytes;
    template <typename V> size_t add(const std::vector<V> &v) {
        const size_t off = (bytes.size() + 255) & ~size_t{255};
        bytes.resize(off + v.size() * sizeof(V));
        if (!v.empty()) {
            std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(V));
        }
        return off;
    }
};

// Appends src's circuits to dst, rebasing every index and offset.
template <typename T> void merge_plan(BatchPlan<T> &dst, BatchPlan<T> &&src) {
    const int32_t pass_off = static_cast<int32_t>(dst.passes.size());
    const int32_t bundle_off = static_cast<int32_t>(dst.bundles.size());
    const int32_t term_off = static_cast<int32_t>(dst.terms.size());
    const int32_t slot_off = static_cast<int32_t>(dst.slot_cols.size());
    dst.ops.resize((dst.ops.size() + 15) & ~size_t{15}, 0);
    const size_t ops_off = dst.ops.size();
    if (ops_off + src.ops.size() > static_cast<size_t>(INT32_MAX)) {
        raise(PLK_ERR_UNSUPPORTED, "operation table exceeds 2 GiB");
    }
    for (auto b : src.bundles) {
        b.first += static_cast<int32_t>(ops_off);
        dst.bundles.push_back(b);
    }
    for (auto p : src.passes) {
        p.bundle_first += bundle_off;
        p.slot_first += slot_off;
        dst.passes.push_back(p);
    }
    for (auto c : src.circuits) {
        c.fwd_first += pass_off;
        c.bwd_first += pass_off;
        c.term_first += term_off;
        c.jac_offset += dst.jac_len;
        dst.circuits.push_back(c);
    }
    dst.ops.insert(dst.ops.end(), src.ops.begin(), src.ops.end());
    dst.terms.insert(dst.terms.end(), src.terms.begin(), src.terms.end());
    dst.slot_cols.insert(dst.slot_cols.end(), src.slot_cols.begin(), src.slot_cols.end());
    dst.slot_coefs.insert(dst.slot_coefs.end(), src.slot_coefs.begin(), src.slot_coefs.end());
    dst.cost.insert(dst.cost.end(), src.cost.begin(), src.cost.end());
    for (auto p : src.patches) {
        p.off += static_cast<int64_t>(ops_off);
        p.param += static_cast<int32_t>(dst.nparams);
        dst.patches.push_back(p);
    }
    dst.jac_len += src.jac_len;
    dst.nparams += src.nparams;
}

// Runs fn(t, lo, hi) for nthreads contiguous ranges of [0, count) on host
// threads (inline when there is one).
template <typename F> void parallel_ranges(size_t count, size_t nthreads, F fn) {
    nthreads = std::max<size_t>(1, std::min(nthreads, count));
    if (nthreads == 1) {
        fn(size_t{0}, size_t{0}, count);
        return;
    }
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nthreads; t++) {
        pool.emplace_back(fn, t, count * t / nthreads, count * (t + 1) / nthreads);
    }
    for (auto &th : pool) {
        th.join();
    }
}

size_t plan_threads(size_t count) {
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    return std::min<size_t>(
        {static_cast<size_t>(std::max(1, env_int("PLK_PLAN_THREADS", static_cast<int>(hw)))),
         (count + 31) / 32, size_t{64}});
}

void hash_obs(Hasher &H, const ObsNode &o) {
    H.add(static_cast<uint64_t>(o.kind));
    H.add(o.name);
    H.add(static_cast<uint64_t>(o.wires.size()));
    for (auto w : o.wires) {
        H.add(static_cast<uint64_t>(w));
    }
    H.add(static_cast<uint64_t>(o.coeffs.size()));
    for (double c : o.coeffs) {
        H.add(c);
    }
    H.add(static_cast<uint64_t>(o.matrix.size()));
    for (const cd &m : o.matrix) {
        H.add(m.real());
        H.add(m.imag());
    }
    H.add(static_cast<uint64_t>(o.terms.size()));
    for (const auto &t : o.terms) {
        hash_obs(H, *t);
    }
}

// Everything a batch plan depends on except the angles of patched gates:
// qubits, op names / wires / inverse flags, the values of every other
// parameter and matrix, the trainable set and the observable.
uint64_t circuit_key(const BatchInput &in) {
    Hasher H;
    const OpsData &ops = *in.ops;
    H.add(static_cast<uint64_t>(in.n));
    H.add(static_cast<uint64_t>(ops.names.size()));
    for (size_t i = 0; i < ops.names.size(); i++) {
        H.add(ops.names[i]);
        H.add(static_cast<uint64_t>(ops.wires[i].size() | (ops.inverses[i] ? size_t{1} << 40 : 0)));
        for (auto w : ops.wires[i]) {
            H.add(static_cast<uint64_t>(w));
        }
        H.add(static_cast<uint64_t>(ops.params[i].size()));
        if (patch_gate_of(ops, i) == BPG_NONE) {
            for (double v : ops.params[i]) {
                H.add(v);
            }
        }
        H.add(static_cast<uint64_t>(ops.matrices[i].size()));
        for (const cd &m : ops.matrices[i]) {
            H.add(m.real());
            H.add(m.imag());
        }
    }
    H.add(static_cast<uint64_t>(in.tp.size()));
    for (auto t : in.tp) {
        H.add(static_cast<uint64_t>(t));
    }
    hash_obs(H, *in.obs);
    return H.h;
}

// The last batch plan of one precision, resident on its device: tables,
// patch list, parameter slots, scratch and result buffers in one allocation.
// A call whose circuits have the same keys reuses it and uploads only the
// parameters; any other call replaces it.
template <typename T> struct BatchCache {
    bool valid = false;
    int device = -1;
    int tile_bits = 0;
    std::vector<uint64_t> keys;
    std::vector<int64_t> param_base; // first global parameter of each circuit
    int64_t nparams = 0;
    int64_t npatches = 0;
    void *dbuf = nullptr;
    void *ops = nullptr;
    const BatchPatch *patches = nullptr;
    double *params = nullptr;
    BatchArgs A{};

    void release() {
        valid = false;
        if (dbuf != nullptr) {
            PLK_HIP(hipStreamSynchronize(device_stream()));
            PLK_HIP(hipFree(dbuf));
            dbuf = nullptr;
        }
    }
    hipStream_t device_stream() const { return ::plk::device(device).stream; }
};

template <typename T> BatchCache<T> &batch_cache() {
    static BatchCache<T> c; // released with the process
    return c;
}

// Plans every circuit and makes the result resident in C.
// Every validation error is raised while planning, before any device call.
template <typename T>
void build_batch_cache(BatchCache<T> &C, int device_id, const std::vector<BatchInput> &in,
                       int tile_bits) {
    const int dtype = sizeof(T) == 4 ? PLK_C64 : PLK_C128;
    const size_t nthreads = plan_threads(in.size());
    std::unique_ptr<HostScope> phase = std::make_unique<HostScope>("host_batch_plan");
    std::vector<BatchPlan<T>> parts(nthreads);
    std::vector<std::pair<size_t, Error>> failures(nthreads, {SIZE_MAX, Error{PLK_OK, ""}});
    parallel_ranges(in.size(), nthreads, [&](size_t t, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) {
            try {
                plan_batch_circuit<T>(parts[t], in[i], tile_bits, low_bits(dtype));
            } catch (const Error &e) {
                failures[t] = {i, e};
                return;
            } catch (const std::exception &e) {
                failures[t] = {i, Error{PLK_ERR_INVALID, e.what()}};
                return;
            }
        }
    });
    for (const auto &f : failures) {
        if (f.first != SIZE_MAX) {
            raise(f.second.code, "circuit " + std::to_string(f.first) + ": " + f.second.msg);
        }
    }
    phase = std::make_unique<HostScope>("host_batch_merge");
    BatchPlan<T> bp = std::move(parts[0]);
    {
        // One allocation per table instead of a regrowth per part.
        size_t nc = 0, np = 0, nb = 0, no = 0, nt = 0, ns = 0, npt = 0;
        for (const auto &pt : parts) {
            nc += pt.circuits.size();
            np += pt.passes.size();
            nb += pt.bundles.size();
            no += pt.ops.size() + 16;
            nt += pt.terms.size();
            ns += pt.slot_cols.size();
            npt += pt.patches.size();
        }
        bp.circuits.reserve(nc);
        bp.cost.reserve(nc);
        bp.passes.reserve(np);
        bp.bundles.reserve(nb);
        bp.ops.reserve(no);
        bp.terms.reserve(nt);
        bp.slot_cols.reserve(ns);
        bp.slot_coefs.reserve(ns);
        bp.patches.reserve(npt);
    }
    for (size_t t = 1; t < parts.size(); t++) {
        merge_plan(bp, std::move(parts[t]));
        BatchPlan<T>().swap_into(parts[t]);
    }
    phase = std::make_unique<HostScope>("host_batch_pack");
    Device &dev = device(device_id);

    // Longest-processing-time-first packing of circuits onto workgroups.
    const int per_cu = batch_resident_blocks(dtype, tile_bits);
    const int grid = static_cast<int>(
        std::min<size_t>(in.size(), static_cast<size_t>(dev.num_cus) * per_cu));
    std::vector<int32_t> by_cost(in.size());
    for (size_t i = 0; i < in.size(); i++) {
        by_cost[i] = static_cast<int32_t>(i);
    }
    std::stable_sort(by_cost.begin(), by_cost.end(),
                     [&](int32_t a, int32_t b) { return bp.cost[a] > bp.cost[b]; });
    using Load = std::pair<double, int>;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (int b = 0; b < grid; b++) {
        heap.push({0.0, b});
    }
    std::vector<std::vector<int32_t>> lists(grid);
    for (int32_t ci : by_cost) {
        Load l = heap.top();
        heap.pop();
        lists[l.second].push_back(ci);
        l.first += bp.cost[ci];
        heap.push(l);
    }
    std::vector<int32_t> order, block_first{0};
    std::vector<int64_t> block_scratch;
    int64_t scratch_amps = 0;
    for (int b = 0; b < grid; b++) {
        int64_t need = 0;
        for (int32_t ci : lists[b]) {
            order.push_back(ci);
            if (!bp.circuits[ci].resident) {
                need = std::max<int64_t>(need, int64_t{2} << bp.circuits[ci].n);
            }
        }
        block_first.push_back(static_cast<int32_t>(order.size()));
        block_scratch.push_back(scratch_amps);
        scratch_amps += need;
    }

    Blob blob;
    const size_t o_circ = blob.add(bp.circuits);
    const size_t o_order = blob.add(order);
    const size_t o_bfirst = blob.add(block_first);
    const size_t o_bscr = blob.add(block_scratch);
    const size_t o_pass = blob.add(bp.passes);
    const size_t o_bund = blob.add(bp.bundles);
    const size_t o_ops = blob.add(bp.ops);
    const size_t o_terms = blob.add(bp.terms);
    const size_t o_cols = blob.add(bp.slot_cols);
    const size_t o_coefs = blob.add(bp.slot_coefs);
    const size_t o_patch = blob.add(bp.patches);
    auto up = [](size_t x) { return (x + 255) & ~size_t{255}; };
    const size_t amp = 2 * sizeof(T);
    const size_t o_params = up(blob.bytes.size());
    const size_t o_scratch = o_params + up(static_cast<size_t>(bp.nparams) * sizeof(double));
    const size_t o_exp = o_scratch + up(static_cast<size_t>(scratch_amps) * amp);
    const size_t o_jac = o_exp + up(in.size() * sizeof(double));
    const size_t total = o_jac + static_cast<size_t>(bp.jac_len) * sizeof(double) + 256;

    phase = std::make_unique<HostScope>("host_batch_upload");
    C.release();
    C.device = dev.id;
    C.dbuf = dev_malloc(dev.id, total, "batch plan");
    auto *w = static_cast<unsigned char *>(C.dbuf);
    PLK_HIP(hipMemcpy(w, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice));
    record_kernel("host_batch_table_bytes", 0.f, static_cast<double>(blob.bytes.size()));

    BatchArgs &A = C.A;
    A = BatchArgs{};
    A.circuits = reinterpret_cast<const BatchCircuit *>(w + o_circ);
    A.order = reinterpret_cast<const int32_t *>(w + o_order);
    A.block_first = reinterpret_cast<const int32_t *>(w + o_bfirst);
    A.block_scratch = reinterpret_cast<const int64_t *>(w + o_bscr);
    A.passes = reinterpret_cast<const BatchPass *>(w + o_pass);
    A.bundles = reinterpret_cast<const TileBundle *>(w + o_bund);
    A.ops = w + o_ops;
    A.terms = reinterpret_cast<const BatchTerm *>(w + o_terms);
    A.slot_cols = reinterpret_cast<const int32_t *>(w + o_cols);
    A.slot_coefs = reinterpret_cast<const double *>(w + o_coefs);
    A.scratch = w + o_scratch;
    A.expvals = reinterpret_cast<double *>(w + o_exp);
    A.jac = reinterpret_cast<double *>(w + o_jac);
    A.jac_len = bp.jac_len;
    A.tile_bits = tile_bits;
    A.grid = grid;
    C.ops = w + o_ops;
    C.patches = reinterpret_cast<const BatchPatch *>(w + o_patch);
    C.npatches = static_cast<int64_t>(bp.patches.size());
    C.params = reinterpret_cast<double *>(w + o_params);
    C.nparams = bp.nparams;
    C.tile_bits = tile_bits;
    C.param_base.assign(in.size(), 0);
    int64_t base = 0;
    for (size_t i = 0; i < in.size(); i++) {
        C.param_base[i] = base;
        base += static_cast<int64_t>(in[i].ops->num_par_ops());
    }
    C.valid = true;
}

template <typename T>
void adjoint_batched_t(int device_id, const std::vector<BatchInput> &in, double *expvals,
                       double *jac) {
    const int dtype = sizeof(T) == 4 ? PLK_C64 : PLK_C128;
    const int tile_bits = batch_tile_bits(dtype);
    if (in.empty()) {
        return;
    }
    const size_t nthreads = plan_threads(in.size());
    std::unique_ptr<HostScope> phase = std::make_unique<HostScope>("host_batch_key");
    std::vector<uint64_t> keys(in.size());
    parallel_ranges(in.size(), nthreads, [&](size_t, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) {
            keys[i] = circuit_key(in[i]);
        }
    });
    phase.reset();
    BatchCache<T> &C = batch_cache<T>();
    if (!(C.valid && C.device == device_id && C.tile_bits == tile_bits && C.keys == keys)) {
        C.valid = false;
        build_batch_cache<T>(C, device_id, in, tile_bits);
        C.keys = std::move(keys);
    }
    Device &dev = device(device_id);

    // Angles of this call, in the plan's parameter order, then the patch
    // kernel rewrites every patched U payload on the device.
    phase = std::make_unique<HostScope>("host_batch_params");
    std::vector<double> params(static_cast<size_t>(C.nparams));
    parallel_ranges(in.size(), nthreads, [&](size_t, size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; c++) {
            const OpsData &ops = *in[c].ops;
            int64_t k = C.param_base[c];
            for (size_t i = 0; i < ops.params.size(); i++) {
                if (!ops.params[i].empty()) {
                    params[static_cast<size_t>(k++)] = ops.params[i][0];
                }
            }
        }
    });
    PLK_HIP(hipMemcpyAsync(C.params, params.data(), params.size() * sizeof(double),
                           hipMemcpyHostToDevice, dev.stream));
    phase = std::make_unique<HostScope>("host_batch_launch");
    launch_batch_patch(dev, dtype, C.ops, C.patches, C.npatches, C.params);
    launch_batch(dev, dtype, C.A);
    phase = std::make_unique<HostScope>("host_batch_results");
    PLK_HIP(hipMemcpyAsync(expvals, C.A.expvals, in.size() * sizeof(double),
                           hipMemcpyDeviceToHost, dev.stream));
    if (C.A.jac_len > 0) {
        PLK_HIP(hipMemcpyAsync(jac, C.A.jac, static_cast<size_t>(C.A.jac_len) * sizeof(double),
                               hipMemcpyDeviceToHost, dev.stream));
    }
    // Also keeps `params` alive until its upload has been read.
    PLK_HIP(hipStreamSynchronize(dev.stream));
}

std::vector<size_t> to_vec(const size_t *w, size_t n) {
    return std::vector<size_t>(w, w + n);
}

std::vector<cd> to_cvec(const double *m, size_t len) {
    std::vector<cd> v(len);
    for (size_t i = 0; i < len; i++) {
        v[i] = cd{m[2 * i], m[2 * i + 1]};
    }
    return v;
}

State &S(plk_state *p) {
    if (!p) {
        raise(PLK_ERR_INVALID, "null state handle");
    }
    return p->s;
}

const State &S(const plk_state *p) {
    if (!p) {
        raise(PLK_ERR_INVALID, "null state handle");
    }
    return p->s;
}

} // namespace

} // namespace plk

struct plk_obs {
    std::shared_ptr<const plk::ObsNode> node;
};

struct plk_ops {
    plk::OpsData data;
};

using namespace plk;

#define PLK_TRY(body)                                                          \
    try {                                                                      \
        std::lock_guard<std::recursive_mutex> plk_engine_lock_(engine_mutex()); \
        begin_call();                                                          \
        body;                                                                  \
        g_last_error.clear();                                                  \
        return PLK_OK;                                                         \
    } catch (const plk::Error &e) {                                            \
        g_last_error = e.msg;                                                  \
        return e.code;                                                         \
    } catch (const std::bad_alloc &) {                                         \
        g_last_error = "host allocation failed";                               \
        return PLK_ERR_OOM;                                                    \
    } catch (const std::exception &e) {                                        \
        g_last_error = e.what();                                               \
        return PLK_ERR_INVALID;                                                \
    }

static const ObsNode &O(const plk_obs *o) {
    if (!o || !o->node) {
        raise(PLK_ERR_INVALID, "null observable handle");
    }
    return *o->node;
}

extern "C" {

const char *plk_last_error(void) { return g_last_error.c_str(); }

int plk_device_count(int *out) {
    PLK_TRY({
        int c = 0;
        PLK_HIP(hipGetDeviceCount(&c));
        *out = c;
    })
}

int plk_print_configuration(char *buf, size_t len) {
    PLK_TRY({
        std::ostringstream os;
        int count = 0;
        PLK_HIP(hipGetDeviceCount(&count));
        os << "plk_amd state-vector engine\n";
        os << "  HIP runtime devices: " << count << "\n";
        for (int d = 0; d < count; d++) {
            hipDeviceProp_t p;
            PLK_HIP(hipGetDeviceProperties(&p, d));
            os << "  device " << d << ": " << p.name << " arch=" << p.gcnArchName
               << " CUs=" << p.multiProcessorCount
               << " HBM=" << (p.totalGlobalMem >> 20) << " MiB"
               << " LDS/block=" << p.sharedMemPerBlock << " B"
               << " wavefront=" << p.warpSize << "\n";
        }
        os << "  build: gfx950 HIP kernels, complex64/complex128 states, FP64 reductions\n";
        os << "  fused segments: up to " << SEG_KMAX << " wires, " << SEG_PMAX
           << " trainable ops and " << SEG_NJMAX << " observable states per HBM pass\n";
        const std::string s = os.str();
        if (len > 0) {
            std::strncpy(buf, s.c_str(), len - 1);
            buf[len - 1] = '\0';
        }
    })
}

int plk_state_create(int num_qubits, int dtype, int device_id, plk_state **out) {
    PLK_TRY({
        if (num_qubits < 0 || num_qubits > 40) {
            raise(PLK_ERR_INVALID, "number of qubits out of range");
        }
        if (dtype != PLK_C64 && dtype != PLK_C128) {
            raise(PLK_ERR_INVALID, "dtype must be PLK_C64 or PLK_C128");
        }
        Device &dev = device(device_id);
        auto h = std::make_unique<plk_state>();
        h->s.n = num_qubits;
        h->s.dtype = dtype;
        h->s.len = size_t{1} << num_qubits;
        h->s.device = device_id;
        h->s.d = dev_malloc(device_id, h->s.bytes(), "state vector");
        launch_set_basis(dev, h->s, 0);
        PLK_HIP(hipStreamSynchronize(dev.stream));
        *out = h.release();
    })
}

int plk_state_destroy(plk_state *sv) {
    PLK_TRY({
        if (sv) {
            Device &dev = device(sv->s.device);
            PLK_HIP(hipStreamSynchronize(dev.stream));
            PLK_HIP(hipFree(sv->s.d));
            delete sv;
        }
    })
}

int plk_state_num_qubits(const plk_state *sv, int *out) { PLK_TRY({ *out = S(sv).n; }) }

int plk_state_length(const plk_state *sv, size_t *out) { PLK_TRY({ *out = S(sv).len; }) }

int plk_state_dtype(const plk_state *sv, int *out) { PLK_TRY({ *out = S(sv).dtype; }) }

int plk_state_device_ptr(plk_state *sv, void **out) { PLK_TRY({ *out = S(sv).d; }) }

int plk_state_upload(plk_state *sv, const void *host, size_t length) {
    PLK_TRY({
        State &s = S(sv);
        if (length != s.len) {
            raise(PLK_ERR_INVALID, "HostToDevice: length " + std::to_string(length) +
                                       " does not match the state length " +
                                       std::to_string(s.len));
        }
        Device &dev = dev_of(s);
        PLK_HIP(hipMemcpyAsync(s.d, host, s.bytes(), hipMemcpyHostToDevice, dev.stream));
        PLK_HIP(hipStreamSynchronize(dev.stream));
    })
}

int plk_state_download(const plk_state *sv, void *host, size_t length) {
    PLK_TRY({
        const State &s = S(sv);
        if (length != s.len) {
            raise(PLK_ERR_INVALID, "DeviceToHost: length " + std::to_string(length) +
                                       " does not match the state length " +
                                       std::to_string(s.len));
        }
        Device &dev = dev_of(s);
        PLK_HIP(hipMemcpyAsync(host, s.d, s.bytes(), hipMemcpyDeviceToHost, dev.stream));
        PLK_HIP(hipStreamSynchronize(dev.stream));
    })
}

int plk_state_reset(plk_state *sv) {
    PLK_TRY({
        State &s = S(sv);
        launch_set_basis(dev_of(s), s, 0);
    })
}

int plk_state_set_basis(plk_state *sv, size_t index) {
    PLK_TRY({
        State &s = S(sv);
        if (index >= s.len) {
            raise(PLK_ERR_INVALID, "basis-state index out of range");
        }
        launch_set_basis(dev_of(s), s, index);
    })
}

int plk_state_set_values(plk_state *sv, const size_t *indices, const void *values,
                         size_t count) {
    PLK_TRY({
        State &s = S(sv);
        Device &dev = dev_of(s);
        for (size_t i = 0; i < count; i++) {
            if (indices[i] >= s.len) {
                raise(PLK_ERR_INVALID, "setStateVector: index out of range");
            }
        }
        launch_fill_zero(dev, s);
        if (count) {
            void *blob = nullptr;
            const size_t ib = count * sizeof(size_t);
            const size_t vo = (ib + 255) / 256 * 256;
            PLK_HIP(hipMalloc(&blob, vo + count * s.elem_bytes()));
            PLK_HIP(hipMemcpy(blob, indices, ib, hipMemcpyHostToDevice));
            PLK_HIP(hipMemcpy(static_cast<char *>(blob) + vo, values, count * s.elem_bytes(),
                              hipMemcpyHostToDevice));
            launch_scatter(dev, s, static_cast<const size_t *>(blob),
                           static_cast<char *>(blob) + vo, count);
            PLK_HIP(hipStreamSynchronize(dev.stream));
            PLK_HIP(hipFree(blob));
        }
    })
}

int plk_state_copy(plk_state *dst, const plk_state *src) {
    PLK_TRY({
        State &d = S(dst);
        const State &s = S(src);
        if (d.device != s.device) {
            raise(PLK_ERR_INVALID, "state copy across devices is not supported");
        }
        launch_copy(dev_of(d), d, s);
    })
}

int plk_state_synchronize(const plk_state *sv) {
    PLK_TRY({ PLK_HIP(hipStreamSynchronize(dev_of(S(sv)).stream)); })
}

int plk_apply(plk_state *sv, const char *name, const size_t *wires, size_t num_wires,
              int inverse, const double *params, size_t num_params) {
    PLK_TRY({
        State &s = S(sv);
        const std::string nm(name ? name : "");
        if (!is_named_gate(nm)) {
            raise(PLK_ERR_UNSUPPORTED, "operation " + nm + " has no named kernel");
        }
        std::vector<double> p(params, params + num_params);
        apply_single(dev_of(s), s,
                     compile_op(nm, to_vec(wires, num_wires), inverse != 0, p, {}, s.n));
    })
}

int plk_apply_matrix(plk_state *sv, const double *matrix, const size_t *wires,
                     size_t num_wires, int inverse) {
    PLK_TRY({
        State &s = S(sv);
        const size_t dim = size_t{1} << num_wires;
        apply_single(dev_of(s), s,
                     compile_op("Matrix", to_vec(wires, num_wires), inverse != 0, {},
                                to_cvec(matrix, dim * dim), s.n));
    })
}

int plk_apply_ops(plk_state *sv, const plk_ops *ops) {
    PLK_TRY({
        State &s = S(sv);
        if (!ops) {
            raise(PLK_ERR_INVALID, "null ops handle");
        }
        Device &dev = dev_of(s);
        run_forward(dev, s, forward_plan(ops->data, s.n, s.dtype));
    })
}

int plk_apply_ops_blocks(plk_state *sv, const plk_ops *ops, size_t *num_blocks) {
    PLK_TRY({
        State &s = S(sv);
        if (!ops) {
            raise(PLK_ERR_INVALID, "null ops handle");
        }
        if (s.n < 9) {
            raise(PLK_ERR_UNSUPPORTED, "5-qubit MFMA blocks need at least 9 qubits");
        }
        const OpsData &d = ops->data;
        std::vector<CompiledOp> cops;
        cops.reserve(d.names.size());
        for (size_t i = 0; i < d.names.size(); i++) {
            cops.push_back(compile_op(d.names[i], d.wires[i], d.inverses[i], d.params[i],
                                      d.matrices[i], s.n));
        }
        const std::vector<FusedBlock> blocks = fuse_blocks(cops, s.n);
        apply_blocks(dev_of(s), s, blocks);
        if (num_blocks) {
            *num_blocks = blocks.size();
        }
    })
}

int plk_apply_generator(plk_state *sv, const char *name, const size_t *wires,
                        size_t num_wires, int adjoint, double *scale) {
    PLK_TRY({
        (void)adjoint;
        State &s = S(sv);
        CompiledGen g = compile_generator(name ? name : "", to_vec(wires, num_wires), s.n);
        apply_generator_dev(dev_of(s), s, g);
        *scale = g.scale;
    })
}

int plk_expval_named(const plk_state *sv, const char *name, const size_t *wires,
                     size_t num_wires, double *out) {
    PLK_TRY({
        const State &s = S(sv);
        ObsNode o;
        o.kind = ObsNode::Named;
        o.name = name ? name : "";
        o.wires = to_vec(wires, num_wires);
        std::vector<PauliTerm> terms;
        if (!to_pauli(o, s.n, terms)) {
            raise(PLK_ERR_INVALID, "named observable " + o.name + " is not supported");
        }
        *out = pauli_expval(dev_of(s), s, terms);
    })
}

int plk_expval_matrix(const plk_state *sv, const double *matrix, const size_t *wires,
                      size_t num_wires, double *out) {
    PLK_TRY({
        const State &s = S(sv);
        ObsNode o;
        o.kind = ObsNode::Hermitian;
        o.wires = to_vec(wires, num_wires);
        const size_t dim = size_t{1} << num_wires;
        o.matrix = to_cvec(matrix, dim * dim);
        *out = expval_obs(dev_of(s), s, o);
    })
}

int plk_expval_csr(const plk_state *sv, const double *data, const int64_t *indices,
                   const int64_t *indptr, size_t nnz, size_t num_rows, double *out) {
    PLK_TRY({
        const State &s = S(sv);
        CsrView v;
        v.data = data;
        v.indices = indices;
        v.indptr = indptr;
        v.idx32 = false;
        v.nnz = nnz;
        v.rows = num_rows;
        *out = csr_expval_once(dev_of(s), s, v);
    })
}

int plk_expval_csr_i32(const plk_state *sv, const double *data, const int32_t *indices,
                       const int32_t *indptr, size_t nnz, size_t num_rows, double *out) {
    PLK_TRY({
        const State &s = S(sv);
        if (num_rows > size_t{1} << 31 || nnz > static_cast<size_t>(INT32_MAX)) {
            raise(PLK_ERR_INVALID, "CSR matrix too large for 32-bit indices");
        }
        CsrView v;
        v.data = data;
        v.indices = indices;
        v.indptr = indptr;
        v.idx32 = true;
        v.nnz = nnz;
        v.rows = num_rows;
        *out = csr_expval_once(dev_of(s), s, v);
    })
}

int plk_expval_obs(const plk_state *sv, const plk_obs *ob, double *out) {
    PLK_TRY({
        const State &s = S(sv);
        *out = expval_obs(dev_of(s), s, O(ob));
    })
}

int plk_expval_obs_noalloc(const plk_state *sv, const plk_obs *ob, double *out) {
    PLK_TRY({
        const State &s = S(sv);
        std::vector<PauliTerm> terms;
        // Expanded past the limit so that an oversized sum is reported by size.
        if (!to_pauli(O(ob), s.n, terms, 4 * kNoAllocMaxPauliWords)) {
            raise(PLK_ERR_UNSUPPORTED,
                  "plk_expval_obs_noalloc: the observable is not a sum of at most " +
                      std::to_string(kNoAllocMaxPauliWords) +
                      " products of Identity / PauliX / PauliY / PauliZ / Hadamard");
        }
        *out = pauli_expval_noalloc(dev_of(s), s, terms);
    })
}

int plk_inner_product(const plk_state *x, const plk_state *y, double *re, double *im) {
    PLK_TRY({
        const State &a = S(x);
        const State &b = S(y);
        if (a.len != b.len || a.dtype != b.dtype) {
            raise(PLK_ERR_INVALID, "inner product of states with different shapes");
        }
        const State *xs[1] = {&a};
        double ri[2];
        inner_products(dev_of(b), xs, 1, b, ri);
        *re = ri[0];
        *im = ri[1];
    })
}

int plk_obs_named(const char *name, const size_t *wires, size_t num_wires, plk_obs **out) {
    PLK_TRY({
        auto n = std::make_shared<ObsNode>();
        n->kind = ObsNode::Named;
        n->name = name ? name : "";
        n->wires = to_vec(wires, num_wires);
        *out = new plk_obs{n};
    })
}

int plk_obs_hermitian(const double *matrix, const size_t *wires, size_t num_wires,
                      plk_obs **out) {
    PLK_TRY({
        auto n = std::make_shared<ObsNode>();
        n->kind = ObsNode::Hermitian;
        n->wires = to_vec(wires, num_wires);
        const size_t dim = size_t{1} << num_wires;
        n->matrix = to_cvec(matrix, dim * dim);
        *out = new plk_obs{n};
    })
}

int plk_obs_tensor(const plk_obs *const *factors, size_t num_factors, plk_obs **out) {
    PLK_TRY({
        auto n = std::make_shared<ObsNode>();
        n->kind = ObsNode::Tensor;
        std::set<size_t> seen;
        for (size_t i = 0; i < num_factors; i++) {
            const ObsNode &f = O(factors[i]);
            for (auto w : obs_wires(f)) {
                if (!seen.insert(w).second) {
                    raise(PLK_ERR_INVALID, "All wires in observables must be disjoint.");
                }
            }
            n->terms.push_back(factors[i]->node);
        }
        *out = new plk_obs{n};
    })
}

int plk_obs_hamiltonian(const double *coeffs, const plk_obs *const *terms, size_t num_terms,
                        plk_obs **out) {
    PLK_TRY({
        auto n = std::make_shared<ObsNode>();
        n->kind = ObsNode::Hamiltonian;
        for (size_t i = 0; i < num_terms; i++) {
            O(terms[i]);
            n->terms.push_back(terms[i]->node);
            n->coeffs.push_back(coeffs[i]);
        }
        *out = new plk_obs{n};
    })
}

int plk_obs_sparse(const double *data, const int64_t *indices, const int64_t *indptr,
                   size_t nnz, size_t num_rows, const size_t *wires, size_t num_wires,
                   plk_obs **out) {
    PLK_TRY({
        auto n = std::make_shared<ObsNode>();
        n->kind = ObsNode::Sparse;
        n->data.assign(data, data + 2 * nnz);
        n->indices.assign(indices, indices + nnz);
        n->indptr.assign(indptr, indptr + num_rows + 1);
        n->rows = num_rows;
        n->wires = to_vec(wires, num_wires);
        *out = new plk_obs{n};
    })
}

int plk_obs_sparse_i32(const double *data, const int32_t *indices, const int32_t *indptr,
                       size_t nnz, size_t num_rows, const size_t *wires, size_t num_wires,
                       plk_obs **out) {
    PLK_TRY({
        if (num_rows > size_t{1} << 31 || nnz > static_cast<size_t>(INT32_MAX)) {
            raise(PLK_ERR_INVALID, "CSR matrix too large for 32-bit indices");
        }
        auto n = std::make_shared<ObsNode>();
        n->kind = ObsNode::Sparse;
        n->idx32 = true;
        n->data.assign(data, data + 2 * nnz);
        n->indices32.assign(indices, indices + nnz);
        n->indptr32.assign(indptr, indptr + num_rows + 1);
        n->rows = num_rows;
        n->wires = to_vec(wires, num_wires);
        *out = new plk_obs{n};
    })
}

int plk_obs_destroy(plk_obs *ob) { PLK_TRY({ delete ob; }) }

int plk_obs_wires(const plk_obs *ob, size_t *wires, size_t cap, size_t *count) {
    PLK_TRY({
        auto w = obs_wires(O(ob));
        *count = w.size();
        for (size_t i = 0; i < w.size() && i < cap; i++) {
            wires[i] = w[i];
        }
    })
}

int plk_obs_apply(plk_state *sv, const plk_obs *ob) {
    PLK_TRY({
        State &s = S(sv);
        Device &dev = dev_of(s);
        if (const PauliDevice *t = O(ob).pauli_on(dev, s.n)) {
            TmpState tmp(s.n, s.dtype, s.device);
            launch_pauli_apply(dev, tmp.s, s, *t);
            launch_copy(dev, s, tmp.s);
            PLK_HIP(hipStreamSynchronize(dev.stream));
        } else {
            apply_obs_general(dev, s, O(ob));
        }
    })
}

int plk_ops_create(size_t num_ops, const char *const *names, const double *params,
                   const size_t *param_offsets, const size_t *wires,
                   const size_t *wire_offsets, const int *inverses, const double *matrices,
                   const size_t *matrix_offsets, plk_ops **out) {
    PLK_TRY({
        auto h = std::make_unique<plk_ops>();
        OpsData &d = h->data;
        for (size_t i = 0; i < num_ops; i++) {
            d.names.emplace_back(names[i]);
            d.params.emplace_back(params + param_offsets[i], params + param_offsets[i + 1]);
            d.wires.emplace_back(wires + wire_offsets[i], wires + wire_offsets[i + 1]);
            d.inverses.push_back(inverses[i] != 0);
            const size_t ml = matrix_offsets[i + 1] - matrix_offsets[i];
            d.matrices.push_back(ml ? to_cvec(matrices + 2 * matrix_offsets[i], ml)
                                    : std::vector<cd>{});
        }
        *out = h.release();
    })
}

int plk_ops_destroy(plk_ops *ops) { PLK_TRY({ delete ops; }) }

int plk_ops_num_par_ops(const plk_ops *ops, size_t *out) {
    PLK_TRY({
        if (!ops) {
            raise(PLK_ERR_INVALID, "null ops handle");
        }
        *out = ops->data.num_par_ops();
    })
}

int plk_adjoint_jacobian(const plk_state *sv, const plk_obs *const *obs, size_t num_obs,
                         const plk_ops *ops, const size_t *trainable, size_t num_trainable,
                         int apply_operations, double *jac) {
    PLK_TRY({
        const State &s = S(sv);
        if (!ops) {
            raise(PLK_ERR_INVALID, "null ops handle");
        }
        std::vector<std::shared_ptr<const ObsNode>> ob;
        for (size_t j = 0; j < num_obs; j++) {
            O(obs[j]);
            ob.push_back(obs[j]->node);
        }
        std::vector<size_t> tp(trainable, trainable + num_trainable);
        adjoint_jacobian(s, ob, ops->data, tp, apply_operations != 0, jac);
    })
}

int plk_adjoint_batched(int dtype, int device_id, size_t num_circuits, const int *num_qubits,
                        const plk_ops *const *ops, const plk_obs *const *obs,
                        const size_t *const *trainable, const size_t *num_trainable,
                        double *expvals, double *jac) {
    PLK_TRY({
        if (dtype != PLK_C64 && dtype != PLK_C128) {
            raise(PLK_ERR_INVALID, "dtype must be PLK_C64 or PLK_C128");
        }
        if (num_circuits > 0 && (!num_qubits || !ops || !obs || !expvals)) {
            raise(PLK_ERR_INVALID, "null argument");
        }
        if (trainable && !num_trainable) {
            raise(PLK_ERR_INVALID, "trainable lists need their lengths");
        }
        HostScope whole("host_batch_call");
        std::vector<BatchInput> in(num_circuits);
        std::unique_ptr<HostScope> inputs = std::make_unique<HostScope>("host_batch_inputs");
        for (size_t i = 0; i < num_circuits; i++) {
            auto who = [i]() { return "circuit " + std::to_string(i) + ": "; };
            if (num_qubits[i] < 1 || num_qubits[i] > BATCH_NMAX) {
                raise(PLK_ERR_UNSUPPORTED, who() + "batched circuits have 1.." +
                                               std::to_string(BATCH_NMAX) + " qubits");
            }
            if (!ops[i]) {
                raise(PLK_ERR_INVALID, who() + "null ops handle");
            }
            in[i].n = num_qubits[i];
            in[i].ops = &ops[i]->data;
            in[i].obs = &O(obs[i]);
            if (trainable) {
                if (num_trainable[i] > 0 && !trainable[i]) {
                    raise(PLK_ERR_INVALID, who() + "null trainable list");
                }
                in[i].tp.assign(trainable[i], trainable[i] + num_trainable[i]);
            } else {
                in[i].tp.resize(ops[i]->data.num_par_ops());
                for (size_t k = 0; k < in[i].tp.size(); k++) {
                    in[i].tp[k] = k;
                }
            }
            if (in[i].tp.empty()) {
                raise(PLK_ERR_INVALID, who() + "No trainable parameters provided.");
            }
        }
        inputs.reset();
        if (num_circuits > 0 && !jac) {
            raise(PLK_ERR_INVALID, "null jac buffer");
        }
        if (dtype == PLK_C64) {
            adjoint_batched_t<float>(device_id, in, expvals, jac);
        } else {
            adjoint_batched_t<double>(device_id, in, expvals, jac);
        }
    })
}

int plk_probs(const plk_state *sv, const size_t *wires, size_t num_wires, double *out) {
    PLK_TRY({
        const State &s = S(sv);
        if (!out || (num_wires > 0 && !wires)) {
            raise(PLK_ERR_INVALID, "null argument");
        }
        std::vector<size_t> w(wires, wires + num_wires);
        if (w.empty()) {
            for (int q = 0; q < s.n; q++) {
                w.push_back(static_cast<size_t>(q));
            }
        }
        uint64_t seen = 0;
        for (size_t q : w) {
            if (q >= static_cast<size_t>(s.n)) {
                raise(PLK_ERR_INVALID, "wire " + std::to_string(q) + " out of range");
            }
            if (seen >> q & 1) {
                raise(PLK_ERR_INVALID, "repeated wire in probs");
            }
            seen |= uint64_t{1} << q;
        }
        // The reference sums over the sorted wires and then transposes with
        // getTransposedIndexFunctor (MeasuresKokkos.hpp:487-512,
        // MeasuresFunctors.hpp:144-175); its output equals the marginal whose
        // index bits follow the wire order w[pi[pi[k]]], pi = argsort(w).
        std::vector<size_t> pi(w.size());
        for (size_t k = 0; k < pi.size(); k++) {
            pi[k] = k;
        }
        std::sort(pi.begin(), pi.end(), [&](size_t a, size_t b) { return w[a] < w[b]; });
        std::vector<int> g;
        for (size_t k = 0; k < w.size(); k++) {
            g.push_back(s.n - 1 - static_cast<int>(w[pi[pi[k]]]));
        }
        Device &dev = dev_of(s);
        set_device(s.device);
        probs(dev, s, g, out);
    })
}

int plk_generate_samples(const plk_state *sv, size_t num_shots, uint64_t seed, size_t *samples) {
    PLK_TRY({
        const State &s = S(sv);
        if (num_shots > 0 && !samples) {
            raise(PLK_ERR_INVALID, "null samples buffer");
        }
        Device &dev = dev_of(s);
        set_device(s.device);
        sample(dev, s, num_shots, seed, samples);
    })
}

int plk_profiling_enable(int on) { PLK_TRY({ profiling_enable_impl(on != 0); }) }

int plk_profiling_get(const char *kernel, double *launches, double *total_ms,
                      double *bytes) {
    PLK_TRY({
        KernelStats k = profiling_get_impl(kernel ? kernel : "");
        *launches = k.launches;
        *total_ms = k.ms;
        *bytes = k.bytes;
    })
}

int plk_profiling_names(char *buf, size_t len) {
    PLK_TRY({
        const std::string s = profiling_names_impl();
        if (len > 0) {
            std::strncpy(buf, s.c_str(), len - 1);
            buf[len - 1] = '\0';
        }
    })
}

} // extern "C"
