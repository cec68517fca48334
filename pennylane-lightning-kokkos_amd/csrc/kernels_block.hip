// Dense 5-qubit blocks on the matrix cores (complex64 states).
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
// per byte of complex64 state.
#include "plk_internal.hpp"

namespace plk {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kBlockWaves = 4;

struct BlockArgs {
    uint64_t groups;      // 2^(n-5)
    int bits[5];          // ascending block bits
    int pad;
    uint64_t off_t[8];    // offset of local states t = 0..7 (bits b0, b1, b2)
    uint64_t off_kq[4];   // offset of local states 8 kq (bits b3, b4)
};

__device__ __forceinline__ uint64_t insert_zero_bit(uint64_t x, int b) {
    const uint64_t lo = x & ((uint64_t{1} << b) - 1);
    return ((x >> b) << (b + 1)) | lo;
}

__global__ void __launch_bounds__(64 * kBlockWaves)
    block5_kernel(float2 *__restrict__ psi, const float *__restrict__ apack, BlockArgs a) {
    const int l = threadIdx.x & 63;
    const int j = l & 15;
    const int kq = l >> 4;
    // The lane's 64 matrix values: apack[(rb * 16 + ks) * 64 + l].
    float A[64];
#pragma unroll
    for (int e = 0; e < 64; e++) {
        A[e] = apack[e * 64 + l];
    }
    const uint64_t lane_off = a.off_kq[kq];
    const uint64_t tiles = a.groups >> 4;
    const uint64_t wave = uint64_t{blockIdx.x} * kBlockWaves + (threadIdx.x >> 6);
    const uint64_t nwaves = uint64_t{gridDim.x} * kBlockWaves;
    for (uint64_t tile = wave; tile < tiles; tile += nwaves) {
        uint64_t base = tile * 16 + static_cast<uint64_t>(j);
#pragma unroll
        for (int q = 0; q < 5; q++) {
            base = insert_zero_bit(base, a.bits[q]);
        }
        base += lane_off;
        float vr[8], vi[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const float2 v = psi[base + a.off_t[t]];
            vr[t] = v.x;
            vi[t] = v.y;
        }
        f4 c[4];
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            c[rb] = f4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int ks = 0; ks < 16; ks++) {
            const float b = ks < 8 ? vr[ks] : vi[ks - 8];
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                c[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[rb * 16 + ks], b, c[rb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            // Output row 16 rb + 4 kq + r: Re of local 8 kq + 4 rb + r for
            // rb < 2, Im of local 8 kq + 4 (rb - 2) + r for rb >= 2.
            const int rb = t >> 2, r = t & 3;
            psi[base + a.off_t[t]] = float2{c[rb][r], c[rb + 2][r]};
        }
    }
}

} // namespace

int block5_lane_pack(const std::vector<cd> &M, std::vector<float> &out) {
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
    // Output row o = 16 rb + 4 kq + r -> logical row; input row k = 4 ks + kq
    // -> logical column.
    auto out_row = [](int o) {
        const int rb = o >> 4, kq = (o >> 2) & 3, r = o & 3;
        return rb < 2 ? 8 * kq + 4 * rb + r : 32 + 8 * kq + 4 * (rb - 2) + r;
    };
    auto in_row = [](int k) {
        const int ks = k >> 2, kq = k & 3;
        return ks < 8 ? 8 * kq + ks : 32 + 8 * kq + (ks - 8);
    };
    out.assign(64 * 64, 0.f);
    for (int rb = 0; rb < 4; rb++) {
        for (int ks = 0; ks < 16; ks++) {
            for (int l = 0; l < 64; l++) {
                // A operand of lane l: row 16 rb + l % 16, column 4 ks + l / 16.
                const int o = 16 * rb + (l & 15);
                const int k = 4 * ks + (l >> 4);
                out[static_cast<size_t>((rb * 16 + ks) * 64 + l)] =
                    static_cast<float>(R(out_row(o), in_row(k)));
            }
        }
    }
    return 0;
}

void launch_block5(Device &dev, const State &s, const int *bits, const float *d_apack) {
    if (s.dtype != PLK_C64) {
        raise(PLK_ERR_UNSUPPORTED, "5-qubit MFMA blocks need a complex64 state");
    }
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
    hipLaunchKernelGGL(block5_kernel, dim3(grid), dim3(64 * kBlockWaves), 0, dev.stream,
                       static_cast<float2 *>(s.d), d_apack, a);
    PLK_HIP(hipGetLastError());
}

} // namespace plk
