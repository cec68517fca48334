// Dense 5-qubit blocks on the matrix cores: v_mfma_f32_16x16x4f32 for
// complex64 states, v_mfma_f64_16x16x4f64 for complex128 states.
//
// A block is a 32x32 complex matrix M acting on five state bits b0 < ... < b4
// (local index s = sum_t bit(b_t) << t). Written over the reals, w = M v is
// a 64x64 real product [wr; wi] = [[Mr, -Mi], [Mi, Mr]] [vr; vi], and over
// all 2^(n-5) groups of amplitudes that share the other bits it is one GEMM
// of a fixed 64x64 matrix with a 64 x 2^(n-5) matrix of gathered amplitudes.
// Each wave takes 16 groups at a time and runs it as 64
// v_mfma_f32_16x16x4f32 (4 output row blocks x 16 steps of K = 4).
//
// The rows of both operands are permuted so that no data is exchanged
// between lanes: lane l = 16 kq + j (j = group of the tile, kq = l / 16)
// loads the eight amplitudes s = 8 kq + t, t = 0..7, of its group. The MFMA
// B operand of lane l at step ks is row 4 ks + kq of column j, so row
// 4 ks + kq is defined to be Re v[8 kq + ks] (ks < 8) or Im v[8 kq + ks - 8]
// (ks >= 8): every lane feeds the values it loaded. The output block rb of
// the 16x16x4 instruction puts row 4 kq + r of column j in register r of lane
// 16 kq + j, so output row 16 rb + 4 kq + r is defined to be
// Re w[8 kq + 4 rb + r] (rb < 2) or Im w[8 kq + 4 (rb - 2) + r] (rb >= 2):
// every lane ends up holding the eight amplitudes it loaded, and stores them
// in place. The host applies both permutations to the 64x64 real matrix and
// lays it out per lane (64 floats a lane, resident in registers).
//
// Bytes per launch: 2 S (each amplitude read and written once); flops:
// 8 * 32 per amplitude (a complex multiply-add is 8 real flops), so 16 flop
// per byte of complex64 state (8 per byte of complex128 state).
// The FP64 instruction has the same A / B layouts and puts output row
// 4 r + kq (rather than 4 kq + r) in register r of lane 16 kq + j; only the
// host's output-row permutation differs.
#include "plk_internal.hpp"

namespace plk {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kBlockWaves = 4;

struct BlockArgs {
    uint64_t groups;      // 2^(n-5)
    int bits[5];          // ascending block bits
    int pad;
    uint64_t off_t[8];    // offset of local states t = 0..7 (bits b0, b1, b2)
    uint64_t off_kq[4];   // offset of local states 8 kq (bits b3, b4)
};

template <typename T> struct BlockTypes;
template <> struct BlockTypes<float> {
    using C2 = float2;
    using Acc = f4;
    static __device__ __forceinline__ Acc mfma(float a, float b, Acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct BlockTypes<double> {
    using C2 = double2;
    using Acc = d4;
    static __device__ __forceinline__ Acc mfma(double a, double b, Acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ uint64_t insert_zero_bit(uint64_t x, int b) {
    const uint64_t lo = x & ((uint64_t{1} << b) - 1);
    return ((x >> b) << (b + 1)) | lo;
}

__device__ __forceinline__ uint64_t group_base(const BlockArgs &a, uint64_t g) {
#pragma unroll
    for (int q = 0; q < 5; q++) {
        g = insert_zero_bit(g, a.bits[q]);
    }
    return g;
}

// One wave: tiles of 16 groups, round-robin over the waves of the grid; the
// next tile's amplitudes are loaded while the current tile's products run.
template <typename T>
__global__ void __launch_bounds__(64 * kBlockWaves)
    block5_kernel(typename BlockTypes<T>::C2 *__restrict__ psi, const T *__restrict__ apack,
                  BlockArgs a) {
    using BT = BlockTypes<T>;
    using C2 = typename BT::C2;
    const int l = threadIdx.x & 63;
    const int j = l & 15;
    const int kq = l >> 4;
    // The lane's 64 matrix values: apack[(rb * 16 + ks) * 64 + l].
    T A[64];
#pragma unroll
    for (int e = 0; e < 64; e++) {
        A[e] = apack[e * 64 + l];
    }
    const uint64_t lane_off = a.off_kq[kq];
    const uint64_t tiles = a.groups >> 4;
    const uint64_t wave = uint64_t{blockIdx.x} * kBlockWaves + (threadIdx.x >> 6);
    const uint64_t nwaves = uint64_t{gridDim.x} * kBlockWaves;
    uint64_t tile = wave;
    if (tile >= tiles) {
        return;
    }
    uint64_t base = group_base(a, tile * 16 + static_cast<uint64_t>(j)) + lane_off;
    C2 v[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        v[t] = psi[base + a.off_t[t]];
    }
    while (true) {
        const uint64_t next = tile + nwaves;
        const bool more = next < tiles;
        const uint64_t nbase =
            more ? group_base(a, next * 16 + static_cast<uint64_t>(j)) + lane_off : base;
        C2 nv[8];
        if (more) {
#pragma unroll
            for (int t = 0; t < 8; t++) {
                nv[t] = psi[nbase + a.off_t[t]];
            }
        }
        typename BT::Acc c[4];
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            c[rb] = typename BT::Acc{0, 0, 0, 0};
        }
#pragma unroll
        for (int ks = 0; ks < 16; ks++) {
            const T b = ks < 8 ? v[ks].x : v[ks - 8].y;
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                c[rb] = BT::mfma(A[rb * 16 + ks], b, c[rb]);
            }
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            // Register r of output block rb holds Re (rb < 2) or Im (rb >= 2)
            // of local 8 kq + 4 (rb % 2) + r.
            const int rb = t >> 2, r = t & 3;
            psi[base + a.off_t[t]] = C2{c[rb][r], c[rb + 2][r]};
        }
        if (!more) {
            break;
        }
        tile = next;
        base = nbase;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            v[t] = nv[t];
        }
    }
}

} // namespace

template <typename T>
void block5_lane_pack_t(const std::vector<cd> &M, std::vector<T> &out) {
    // Real 64x64 form R[i][k] with logical index x < 32: Re of local x,
    // x >= 32: Im of local x - 32.
    auto R = [&](int i, int k) -> double {
        const int ri = i >> 5, si = i & 31, rk = k >> 5, sk = k & 31;
        const cd m = M[static_cast<size_t>(si) * 32 + sk];
        if (ri == rk) {
            return m.real();
        }
        return ri == 1 ? m.imag() : -m.imag();
    };
    // Output row o of output block rb = o / 16 sits in register r of lane
    // 16 kq + j with (FP32) o % 16 = 4 kq + r or (FP64) o % 16 = 4 r + kq; it
    // is Re / Im of local 8 kq + 4 (rb % 2) + r. Input row k = 4 ks + kq is
    // Re / Im of local 8 kq + ks % 8.
    auto out_row = [](int o) {
        const int rb = o >> 4;
        const int kq = sizeof(T) == 4 ? (o >> 2) & 3 : o & 3;
        const int r = sizeof(T) == 4 ? o & 3 : (o >> 2) & 3;
        return rb < 2 ? 8 * kq + 4 * rb + r : 32 + 8 * kq + 4 * (rb - 2) + r;
    };
    auto in_row = [](int k) {
        const int ks = k >> 2, kq = k & 3;
        return ks < 8 ? 8 * kq + ks : 32 + 8 * kq + (ks - 8);
    };
    out.assign(64 * 64, T(0));
    for (int rb = 0; rb < 4; rb++) {
        for (int ks = 0; ks < 16; ks++) {
            for (int l = 0; l < 64; l++) {
                // A operand of lane l: row 16 rb + l % 16, column 4 ks + l / 16.
                const int o = 16 * rb + (l & 15);
                const int k = 4 * ks + (l >> 4);
                out[static_cast<size_t>((rb * 16 + ks) * 64 + l)] =
                    static_cast<T>(R(out_row(o), in_row(k)));
            }
        }
    }
}

int block5_lane_pack(const std::vector<cd> &M, std::vector<float> &out) {
    block5_lane_pack_t<float>(M, out);
    return 0;
}

int block5_lane_pack(const std::vector<cd> &M, std::vector<double> &out) {
    block5_lane_pack_t<double>(M, out);
    return 0;
}

void launch_block5(Device &dev, const State &s, const int *bits, const void *d_apack) {
    if (s.n < 9) {
        raise(PLK_ERR_UNSUPPORTED, "5-qubit MFMA blocks need at least 9 qubits");
    }
    BlockArgs a{};
    a.groups = uint64_t{1} << (s.n - 5);
    for (int q = 0; q < 5; q++) {
        a.bits[q] = bits[q];
        if (bits[q] < 0 || bits[q] >= s.n || (q > 0 && bits[q] <= bits[q - 1])) {
            raise(PLK_ERR_INVALID, "block bits must be ascending and inside the state");
        }
    }
    for (int t = 0; t < 8; t++) {
        uint64_t o = 0;
        for (int q = 0; q < 3; q++) {
            if ((t >> q) & 1) {
                o |= uint64_t{1} << bits[q];
            }
        }
        a.off_t[t] = o;
    }
    for (int kq = 0; kq < 4; kq++) {
        uint64_t o = 0;
        for (int q = 0; q < 2; q++) {
            if ((kq >> q) & 1) {
                o |= uint64_t{1} << bits[3 + q];
            }
        }
        a.off_kq[kq] = o;
    }
    const uint64_t tiles = a.groups >> 4;
    const uint64_t want = (tiles + kBlockWaves - 1) / kBlockWaves;
    const unsigned grid = static_cast<unsigned>(
        std::max<uint64_t>(1, std::min<uint64_t>(want, static_cast<uint64_t>(dev.num_cus) * 8)));
    ProfScope prof("apply_block5", 2.0 * static_cast<double>(s.bytes()), dev.stream);
    if (s.dtype == PLK_C64) {
        hipLaunchKernelGGL(block5_kernel<float>, dim3(grid), dim3(64 * kBlockWaves), 0,
                           dev.stream, static_cast<float2 *>(s.d),
                           static_cast<const float *>(d_apack), a);
    } else {
        hipLaunchKernelGGL(block5_kernel<double>, dim3(grid), dim3(64 * kBlockWaves), 0,
                           dev.stream, static_cast<double2 *>(s.d),
                           static_cast<const double *>(d_apack), a);
    }
    PLK_HIP(hipGetLastError());
}

} // namespace plk
