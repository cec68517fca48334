// Fusion of gate runs into dense 5-qubit blocks for the matrix-core kernel
// (kernels_block.hip). The reference applies every gate as its own kernel
// (StateVectorKokkos.hpp:585-676) and dense matrices through
// multiQubitOpFunctor (GateFunctors.hpp:195-300); here runs of gates on at
// most five wires are multiplied into one 32x32 matrix on the host and each
// block is one pass over the state.
#include <set>

#include "plk_internal.hpp"

namespace plk {

namespace {

constexpr int kBlockWires = 5;
constexpr size_t kFuseLookahead = 4096;

// The op's matrix over its own wires (wires[0] most significant).
std::vector<cd> op_matrix(const CompiledOp &op) {
    const size_t m = op.wires.size();
    const size_t dim = size_t{1} << m;
    if (op.kind == OpKind::Dense) {
        return op.mat;
    }
    std::vector<cd> M(dim * dim, 0.0);
    for (size_t r = 0; r < dim; r++) {
        M[r * dim + r] = op.kind == OpKind::Parity
                             ? ((__builtin_popcountll(r) & 1) ? op.odd : op.even)
                             : cd{1.0, 0.0};
    }
    return M;
}

// M <- U M, U acting on the local positions pos[p] of the block (wires[0] of
// the op at pos[0], the most significant bit of U's index).
void left_multiply(std::vector<cd> &M, const std::vector<cd> &U, const std::vector<int> &pos) {
    const int m = static_cast<int>(pos.size());
    const int dim = 1 << m;
    int mask = 0;
    for (int p : pos) {
        mask |= 1 << p;
    }
    auto local = [&](int s) {
        int u = 0;
        for (int p = 0; p < m; p++) {
            u |= ((s >> pos[p]) & 1) << (m - 1 - p);
        }
        return u;
    };
    auto spread = [&](int u, int rest) {
        int s = rest;
        for (int p = 0; p < m; p++) {
            if ((u >> (m - 1 - p)) & 1) {
                s |= 1 << pos[p];
            }
        }
        return s;
    };
    std::vector<cd> out(32 * 32, 0.0);
    for (int r = 0; r < 32; r++) {
        const int ur = local(r);
        const int rest = r & ~mask;
        for (int uc = 0; uc < dim; uc++) {
            const cd u = U[static_cast<size_t>(ur) * dim + uc];
            if (u == cd{0.0, 0.0}) {
                continue;
            }
            const int s = spread(uc, rest);
            for (int c = 0; c < 32; c++) {
                out[static_cast<size_t>(r) * 32 + c] += u * M[static_cast<size_t>(s) * 32 + c];
            }
        }
    }
    M.swap(out);
}

} // namespace

std::vector<FusedBlock> fuse_blocks(const std::vector<CompiledOp> &ops_in, int n) {
    if (n < kBlockWires) {
        raise(PLK_ERR_UNSUPPORTED, "5-qubit blocks need at least 5 qubits");
    }
    std::vector<const CompiledOp *> ops;
    for (const auto &op : ops_in) {
        if (op.kind == OpKind::Identity) {
            continue;
        }
        if (op.wires.size() > static_cast<size_t>(kBlockWires)) {
            raise(PLK_ERR_UNSUPPORTED, "operation " + op.name + " acts on more than five wires");
        }
        ops.push_back(&op);
    }
    std::vector<char> done(ops.size(), 0);
    std::vector<FusedBlock> blocks;
    size_t first = 0;
    while (true) {
        while (first < ops.size() && done[first]) {
            first++;
        }
        if (first == ops.size()) {
            break;
        }
        std::set<size_t> wires(ops[first]->wires.begin(), ops[first]->wires.end());
        std::set<size_t> blocked;
        std::vector<size_t> taken;
        const size_t end = std::min(ops.size(), first + kFuseLookahead);
        for (size_t i = first; i < end; i++) {
            if (done[i]) {
                continue;
            }
            const auto &w = ops[i]->wires;
            bool hits_blocked = false;
            for (auto x : w) {
                hits_blocked = hits_blocked || blocked.count(x) > 0;
            }
            std::set<size_t> grown = wires;
            grown.insert(w.begin(), w.end());
            if (!hits_blocked && grown.size() <= static_cast<size_t>(kBlockWires)) {
                wires.swap(grown);
                taken.push_back(i);
            } else {
                blocked.insert(w.begin(), w.end());
                bool all = true;
                for (auto x : wires) {
                    all = all && blocked.count(x) > 0;
                }
                if (all) {
                    break;
                }
            }
        }
        // Widen to five bits with the lowest bits the block does not hold.
        std::set<int> bits;
        for (auto w : wires) {
            bits.insert(n - 1 - static_cast<int>(w));
        }
        for (int b = 0; static_cast<int>(bits.size()) < kBlockWires; b++) {
            bits.insert(b);
        }
        FusedBlock fb;
        int q = 0;
        for (int b : bits) {
            fb.bits[q++] = b;
        }
        fb.M.assign(32 * 32, 0.0);
        for (int s = 0; s < 32; s++) {
            fb.M[static_cast<size_t>(s) * 33] = 1.0;
        }
        for (size_t i : taken) {
            const CompiledOp &op = *ops[i];
            std::vector<int> pos;
            for (auto w : op.wires) {
                const int b = n - 1 - static_cast<int>(w);
                pos.push_back(static_cast<int>(std::distance(bits.begin(), bits.find(b))));
            }
            left_multiply(fb.M, op_matrix(op), pos);
            done[i] = 1;
        }
        fb.nops = static_cast<int>(taken.size());
        blocks.push_back(std::move(fb));
    }
    return blocks;
}

namespace {

template <typename T>
void apply_blocks_t(Device &dev, const State &s, const std::vector<FusedBlock> &blocks) {
    std::vector<T> all(blocks.size() * 4096);
    std::vector<T> one;
    for (size_t b = 0; b < blocks.size(); b++) {
        block5_lane_pack(blocks[b].M, one);
        std::copy(one.begin(), one.end(), all.begin() + static_cast<std::ptrdiff_t>(b * 4096));
    }
    auto *d = static_cast<T *>(ensure_work(dev, all.size() * sizeof(T)));
    PLK_HIP(hipMemcpyAsync(d, all.data(), all.size() * sizeof(T), hipMemcpyHostToDevice,
                           dev.stream));
    for (size_t b = 0; b < blocks.size(); b++) {
        launch_block5(dev, s, blocks[b].bits, d + b * 4096);
    }
    // The host buffer stays alive until its upload has been read.
    PLK_HIP(hipStreamSynchronize(dev.stream));
}

} // namespace

void apply_blocks(Device &dev, const State &s, const std::vector<FusedBlock> &blocks) {
    if (blocks.empty()) {
        return;
    }
    if (s.dtype == PLK_C64) {
        apply_blocks_t<float>(dev, s, blocks);
    } else {
        apply_blocks_t<double>(dev, s, blocks);
    }
}

} // namespace plk
