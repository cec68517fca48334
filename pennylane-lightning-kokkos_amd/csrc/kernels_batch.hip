// Batched adjoint Jacobians: many independent circuits in one persistent
// launch.
//
// Each 256-thread workgroup runs the circuits the host assigned to it (a
// longest-processing-time-first packing, so the grid drains together and the
// result does not depend on timing). Per circuit:
//   phi = |0...0>; forward passes; h = H phi and <H> (Pauli-word sum);
//   reverse passes: U^dagger on phi and h, Im<h|G phi> of every trainable op
// -- the per-circuit algorithm of AdjointDiffKokkos.hpp:404-478 with one
// observable. A pass holds 2^T amplitudes of phi (and h) in LDS and applies
// bundles of ops on <= 2 tile positions to them. A resident circuit (n <= T)
// never leaves LDS; a larger one lives in the workgroup's slice of a global
// buffer and every pass streams it through LDS tile by tile.
//
// Jacobian sums are FP64 and deterministic: per thread in registers, 16-lane
// DPP row totals, a fixed-order sum over rows, then over tiles in order.
#include "device_common.hpp"

namespace plk {

namespace {

__device__ __forceinline__ uint32_t ins0b(uint32_t k, int b) {
    const uint32_t low = k & ((1u << b) - 1u);
    return ((k >> b) << (b + 1)) | low;
}

// Places the bits of t, lowest first, at the set bits of mask.
__device__ __forceinline__ uint32_t deposit(uint32_t t, uint32_t mask) {
    uint32_t out = 0;
    while (mask) {
        const uint32_t b = mask & (~mask + 1u);
        if (t & 1u) {
            out |= b;
        }
        t >>= 1;
        mask &= mask - 1u;
    }
    return out;
}

// One bundle on a 2^T-amplitude LDS tile: every group of 2^K amplitudes
// (tile positions qa, qb) goes through the bundle's ops in registers.
template <typename T, int K, int NJ>
__device__ __forceinline__ void bundle_tile(Cx<T> *phi, Cx<T> *h, int Tb, const TileBundle &bd,
                                            const unsigned char *ops,
                                            double (&acc)[SEG_PMAX][NJ > 0 ? NJ : 1]) {
    constexpr int D = 1 << K;
    uint32_t o[D];
    if constexpr (K == 1) {
        o[0] = 0;
        o[1] = 1u << bd.qa;
    } else {
        o[0] = 0;
        o[1] = 1u << bd.qb;
        o[2] = 1u << bd.qa;
        o[3] = o[1] | o[2];
    }
    const int groups = (1 << Tb) >> K;
    for (int gi = threadIdx.x; gi < groups; gi += BATCH_BLOCK) {
        uint32_t lb;
        if constexpr (K == 1) {
            lb = ins0b(static_cast<uint32_t>(gi), bd.qa);
        } else {
            lb = ins0b(ins0b(static_cast<uint32_t>(gi), bd.qb), bd.qa);
        }
        Cx<T> f[D];
        Cx<T> hv[NJ > 0 ? NJ : 1][D];
#pragma unroll
        for (int i = 0; i < D; i++) {
            f[i] = phi[lb + o[i]];
            if constexpr (NJ > 0) {
                hv[0][i] = h[lb + o[i]];
            }
        }
        run_ops<T, K, NJ>(f, hv, acc, ops + bd.first, bd.nops);
#pragma unroll
        for (int i = 0; i < D; i++) {
            phi[lb + o[i]] = f[i];
            if constexpr (NJ > 0) {
                h[lb + o[i]] = hv[0][i];
            }
        }
    }
}

struct BatchShared {
    uint32_t lo[64];                            // tile offsets of positions 0..5
    uint32_t hi[64];                            // tile offsets of positions 6..11
    double pacc[BATCH_PMAX];                    // Im<h|G phi> of the pass
    double red[2][BATCH_BLOCK / 16][SEG_PMAX];  // row totals, double-buffered
    double ered[BATCH_BLOCK / 64];
};

// Bundles of one pass over the tile in LDS. NJ = 1 on the reverse walk.
template <typename T, int NJ>
__device__ __forceinline__ void run_bundles(const BatchArgs &A, const BatchPass &P, Cx<T> *lphi,
                                            Cx<T> *lh, BatchShared &sh, int &parity) {
    const unsigned char *ops = static_cast<const unsigned char *>(A.ops);
    const int tid = threadIdx.x;
    for (int b = 0; b < P.bundle_count; b++) {
        const TileBundle bd = A.bundles[P.bundle_first + b];
        double acc[SEG_PMAX][1];
#pragma unroll
        for (int s = 0; s < SEG_PMAX; s++) {
            acc[s][0] = 0.0;
        }
        if (bd.K == 1) {
            bundle_tile<T, 1, NJ>(lphi, lh, P.T, bd, ops, acc);
        } else {
            bundle_tile<T, 2, NJ>(lphi, lh, P.T, bd, ops, acc);
        }
        const bool slots = NJ > 0 && bd.nslots > 0;
        if (slots) {
#pragma unroll
            for (int s = 0; s < SEG_PMAX; s++) {
                if (s < bd.nslots) {
                    const double v = row_total16(acc[s][0]);
                    if ((tid & 15) == 15) {
                        sh.red[parity][tid >> 4][s] = v;
                    }
                }
            }
        }
        __syncthreads();
        if (slots) {
            if (tid < bd.nslots) {
                double tot = 0.0;
                for (int r = 0; r < BATCH_BLOCK / 16; r++) {
                    tot += sh.red[parity][r][tid];
                }
                sh.pacc[bd.slot0 + tid] += tot;
            }
            parity ^= 1;
        }
    }
}

template <typename T, int NJ>
__device__ void run_passes(const BatchArgs &A, const BatchCircuit &c, int first, int count,
                           Cx<T> *lphi, Cx<T> *lh, Cx<T> *gphi, Cx<T> *gh, BatchShared &sh,
                           int &parity) {
    const int tid = threadIdx.x;
    const uint32_t full = (1u << c.n) - 1u;
    for (int p = first; p < first + count; p++) {
        const BatchPass P = A.passes[p];
        if (!c.resident) {
            if (tid < 64) {
                uint32_t v = 0;
                for (int q = 0; q < 6 && q < P.T; q++) {
                    v |= ((static_cast<uint32_t>(tid) >> q) & 1u) << P.tb[q];
                }
                sh.lo[tid] = v;
            } else if (tid < 128) {
                uint32_t v = 0;
                for (int q = 6; q < P.T; q++) {
                    v |= ((static_cast<uint32_t>(tid - 64) >> (q - 6)) & 1u) << P.tb[q];
                }
                sh.hi[tid - 64] = v;
            }
        }
        if constexpr (NJ > 0) {
            if (tid < BATCH_PMAX) {
                sh.pacc[tid] = 0.0;
            }
        }
        __syncthreads();
        const int len = 1 << P.T;
        const int ntiles = c.resident ? 1 : (1 << (c.n - P.T));
        const bool init = (P.flags & BP_INIT) != 0;
        const bool store = (P.flags & BP_NO_STORE) == 0;
        for (int t = 0; t < ntiles; t++) {
            uint32_t base = 0;
            if (!c.resident) {
                base = deposit(static_cast<uint32_t>(t), ~P.tmask & full);
                // Each thread gathers and scatters the same tile positions, so
                // a tile's stores and the next tile's loads need no barrier.
                for (int k = tid; k < len; k += BATCH_BLOCK) {
                    const uint32_t g = base + sh.lo[k & 63] + sh.hi[k >> 6];
                    if (init) {
                        lphi[k] = Cx<T>{static_cast<T>(g == 0 ? 1 : 0), static_cast<T>(0)};
                    } else {
                        lphi[k] = gphi[g];
                    }
                    if constexpr (NJ > 0) {
                        lh[k] = gh[g];
                    }
                }
                __syncthreads();
            }
            run_bundles<T, NJ>(A, P, lphi, lh, sh, parity);
            if (!c.resident && store) {
                for (int k = tid; k < len; k += BATCH_BLOCK) {
                    const uint32_t g = base + sh.lo[k & 63] + sh.hi[k >> 6];
                    gphi[g] = lphi[k];
                    if constexpr (NJ > 0) {
                        gh[g] = lh[k];
                    }
                }
            }
        }
        __syncthreads();
        if constexpr (NJ > 0) {
            if (tid < P.nslots) {
                const int s = P.slot_first + tid;
                A.jac[c.jac_offset + A.slot_cols[s]] = A.slot_coefs[s] * sh.pacc[tid];
            }
        }
    }
}

// h = H phi over the whole state and <phi|H|phi> into expvals[cidx].
template <typename T>
__device__ __forceinline__ void observable(const BatchArgs &A, const BatchCircuit &c, int cidx,
                                           const Cx<T> *phi, Cx<T> *h, BatchShared &sh) {
    const int tid = threadIdx.x;
    const int N = 1 << c.n;
    double e = 0.0;
    for (int i = tid; i < N; i += BATCH_BLOCK) {
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < c.term_count; k++) {
            const BatchTerm t = A.terms[c.term_first + k];
            const Cx<T> p = phi[static_cast<uint32_t>(i) ^ static_cast<uint32_t>(t.x)];
            const double sg = (__popcll(static_cast<uint64_t>(i) & t.zy) & 1) ? -1.0 : 1.0;
            const double pr = static_cast<double>(p.r), pi = static_cast<double>(p.i);
            sr += sg * (t.cr * pr - t.ci * pi);
            si += sg * (t.cr * pi + t.ci * pr);
        }
        const Cx<T> a = phi[i];
        e += static_cast<double>(a.r) * sr + static_cast<double>(a.i) * si;
        h[i] = Cx<T>{static_cast<T>(sr), static_cast<T>(si)};
    }
    e = wave_sum(e);
    if ((tid & 63) == 0) {
        sh.ered[tid >> 6] = e;
    }
    __syncthreads();
    if (tid == 0) {
        double tot = 0.0;
        for (int w = 0; w < BATCH_BLOCK / 64; w++) {
            tot += sh.ered[w];
        }
        A.expvals[cidx] = tot;
    }
}

template <typename T>
__global__ void __launch_bounds__(BATCH_BLOCK) batch_kernel(BatchArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ BatchShared sh;
    const int cap = 1 << A.tile_bits;
    Cx<T> *lphi = reinterpret_cast<Cx<T> *>(smem);
    Cx<T> *lh = lphi + cap;
    const int tid = threadIdx.x;
    Cx<T> *gphi = static_cast<Cx<T> *>(A.scratch) + A.block_scratch[blockIdx.x];
    int parity = 0;
    const int first = A.block_first[blockIdx.x];
    const int last = A.block_first[blockIdx.x + 1];
    for (int ci = first; ci < last; ci++) {
        const int cidx = A.order[ci];
        const BatchCircuit c = A.circuits[cidx];
        const int N = 1 << c.n;
        Cx<T> *gh = gphi + N;
        if (c.resident) {
            for (int i = tid; i < N; i += BATCH_BLOCK) {
                lphi[i] = Cx<T>{static_cast<T>(i == 0 ? 1 : 0), static_cast<T>(0)};
            }
            __syncthreads();
        }
        run_passes<T, 0>(A, c, c.fwd_first, c.fwd_count, lphi, lh, gphi, gh, sh, parity);
        if (c.resident) {
            observable<T>(A, c, cidx, lphi, lh, sh);
        } else {
            observable<T>(A, c, cidx, gphi, gh, sh);
        }
        run_passes<T, 1>(A, c, c.bwd_first, c.bwd_count, lphi, lh, gphi, gh, sh, parity);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) zero_kernel(double *p, int64_t n) {
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += int64_t{gridDim.x} * 256) {
        p[i] = 0.0;
    }
}

template <typename T> void set_batch_lds_limit() {
    static bool attr = false;
    if (!attr) {
        PLK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<T>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(batch_lds_bytes(PLK_C128, BATCH_TMAX))));
        attr = true;
    }
}

template <typename T> void launch_t(Device &dev, const BatchArgs &A) {
    if (A.jac_len > 0) {
        const unsigned g = static_cast<unsigned>(std::min<int64_t>((A.jac_len + 255) / 256, 1024));
        hipLaunchKernelGGL(zero_kernel, dim3(g), dim3(256), 0, dev.stream, A.jac, A.jac_len);
        PLK_HIP(hipGetLastError());
    }
    if (A.grid == 0) {
        return;
    }
    const size_t lds = batch_lds_bytes(sizeof(T) == 4 ? PLK_C64 : PLK_C128, A.tile_bits);
    set_batch_lds_limit<T>();
    ProfScope prof("adjoint_batch", 0.0, dev.stream);
    hipLaunchKernelGGL(batch_kernel<T>, dim3(A.grid), dim3(BATCH_BLOCK), lds, dev.stream, A);
    PLK_HIP(hipGetLastError());
}

} // namespace

size_t batch_lds_bytes(int dtype, int tile_bits) {
    return 2 * (size_t{1} << tile_bits) * (dtype == PLK_C64 ? 8 : 16);
}

int batch_resident_blocks(int dtype, int tile_bits) {
    const size_t lds = batch_lds_bytes(dtype, tile_bits);
    int nb = 0;
    if (dtype == PLK_C64) {
        set_batch_lds_limit<float>();
        PLK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&batch_kernel<float>), BATCH_BLOCK, lds));
    } else {
        set_batch_lds_limit<double>();
        PLK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&batch_kernel<double>), BATCH_BLOCK, lds));
    }
    return std::max(1, std::min(nb, 8));
}

namespace {

// One thread per patch: t = +-params[param], c = cos(t/2), s = sin(t/2).
template <typename T>
__global__ void __launch_bounds__(256)
    batch_patch_kernel(unsigned char *__restrict__ ops, const BatchPatch *__restrict__ patches,
                       int64_t count, const double *__restrict__ params) {
    const int64_t i = int64_t{blockIdx.x} * blockDim.x + threadIdx.x;
    if (i >= count) {
        return;
    }
    const BatchPatch p = patches[i];
    const double t = p.adj ? -params[p.param] : params[p.param];
    T *u = reinterpret_cast<T *>(ops + p.off);
    double s, c;
    if (p.gate == BPG_PHASE) {
        sincos(t, &s, &c);
        u[0] = T(1);
        u[1] = T(0);
        u[2] = static_cast<T>(c);
        u[3] = static_cast<T>(s);
        return;
    }
    sincos(0.5 * t, &s, &c);
    if (p.gate == BPG_RX) {
        // [[c, -i s], [-i s, c]], dense complex row-major.
        u[0] = static_cast<T>(c);
        u[1] = T(0);
        u[2] = T(0);
        u[3] = static_cast<T>(-s);
        u[4] = T(0);
        u[5] = static_cast<T>(-s);
        u[6] = static_cast<T>(c);
        u[7] = T(0);
    } else if (p.gate == BPG_RY) {
        // [[c, -s], [s, c]], real row-major.
        u[0] = static_cast<T>(c);
        u[1] = static_cast<T>(-s);
        u[2] = static_cast<T>(s);
        u[3] = static_cast<T>(c);
    } else if (p.gate == BPG_RZ) {
        // diag(e^{-i t/2}, e^{i t/2}), complex diagonal.
        u[0] = static_cast<T>(c);
        u[1] = static_cast<T>(-s);
        u[2] = static_cast<T>(c);
        u[3] = static_cast<T>(s);
    }
}

} // namespace

void launch_batch_patch(Device &dev, int dtype, void *ops, const BatchPatch *patches,
                        int64_t count, const double *params) {
    if (count <= 0) {
        return;
    }
    const unsigned grid = static_cast<unsigned>((count + 255) / 256);
    ProfScope prof("batch_patch", 0.0, dev.stream);
    if (dtype == PLK_C64) {
        hipLaunchKernelGGL(batch_patch_kernel<float>, dim3(grid), dim3(256), 0, dev.stream,
                           static_cast<unsigned char *>(ops), patches, count, params);
    } else {
        hipLaunchKernelGGL(batch_patch_kernel<double>, dim3(grid), dim3(256), 0, dev.stream,
                           static_cast<unsigned char *>(ops), patches, count, params);
    }
    PLK_HIP(hipGetLastError());
}

void launch_batch(Device &dev, int dtype, const BatchArgs &A) {
    if (dtype == PLK_C64) {
        launch_t<float>(dev, A);
    } else {
        launch_t<double>(dev, A);
    }
}

} // namespace plk
