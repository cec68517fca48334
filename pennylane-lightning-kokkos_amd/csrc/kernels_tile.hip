// LDS tile passes: many fused operations per HBM pass.
//
// A pass owns a set B of T index bits (always including the lowest bits, so
// every tile row is a contiguous run of amplitudes). Each block stages one
// tile -- the 2^T amplitudes with all bits outside B fixed -- of phi and of
// nj observable states in LDS, then applies the pass's bundles in order. A
// bundle is a run of operations on at most two tile bits; each thread loads
// its groups of 2 or 4 amplitudes from LDS into registers, applies the whole
// bundle there (run_ops, shared with the register segment kernel),
// accumulates the FP64 inner products of the bundle's trainable operations
// and writes the group back. Per-wave FP64 partial sums of every trainable
// operation of the pass stay in LDS until the block has processed all its
// tiles, so the reduction is deterministic. One launch reads and writes each
// state once: (1 + w + 2 nj) S bytes for however many operations the pass
// holds.
#include <cstdlib>
#include <map>

#include "device_common.hpp"

namespace plk {

namespace {

__device__ __forceinline__ uint32_t ins0(uint32_t k, int b) {
    const uint32_t low = k & ((1u << b) - 1u);
    return ((k >> b) << (b + 1)) | low;
}

constexpr int kBlockBits = 8;
static_assert((1 << kBlockBits) == TILE_BLOCK, "tile offsets split at the block size");

// Global offset of the tile positions above the block bits for pass k of
// the gather / scatter loop (uniform across the block).
__device__ __forceinline__ uint64_t tile_high_offset(const TileLaunch &L, int k) {
    uint64_t off = 0;
    for (int q = kBlockBits; q < L.T; q++) {
        off |= static_cast<uint64_t>((k >> (q - kBlockBits)) & 1) << L.tb[q];
    }
    return off;
}

template <typename T, int K, int NJ>
__device__ __forceinline__ void run_bundle(Cx<T> *__restrict__ tile, int TS,
                                           const TileBundle &bd, const unsigned char *ops,
                                           double *__restrict__ lacc, int pass_slots) {
    constexpr int D = 1 << K;
    constexpr int NA = NJ > 0 ? NJ : 1;
    double acc[SEG_PMAX][NA];
#pragma unroll
    for (int s = 0; s < SEG_PMAX; s++) {
#pragma unroll
        for (int j = 0; j < NA; j++) {
            acc[s][j] = 0.0;
        }
    }
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
    const int groups = TS >> K;
    for (int gi = threadIdx.x; gi < groups; gi += TILE_BLOCK) {
        uint32_t lb;
        if constexpr (K == 1) {
            lb = ins0(static_cast<uint32_t>(gi), bd.qa);
        } else {
            lb = ins0(ins0(static_cast<uint32_t>(gi), bd.qb), bd.qa);
        }
        Cx<T> f[D];
        Cx<T> h[NA][D];
#pragma unroll
        for (int i = 0; i < D; i++) {
            f[i] = tile[lb + o[i]];
        }
#pragma unroll
        for (int j = 0; j < NJ; j++) {
#pragma unroll
            for (int i = 0; i < D; i++) {
                h[j][i] = tile[(j + 1) * TS + lb + o[i]];
            }
        }
        run_ops<T, K, NJ>(f, h, acc, ops + bd.first, bd.nops);
#pragma unroll
        for (int i = 0; i < D; i++) {
            tile[lb + o[i]] = f[i];
        }
#pragma unroll
        for (int j = 0; j < NJ; j++) {
#pragma unroll
            for (int i = 0; i < D; i++) {
                tile[(j + 1) * TS + lb + o[i]] = h[j][i];
            }
        }
    }
    if constexpr (NJ > 0 && SEG_PMAX * NJ <= 16) {
        if (bd.nslots > 0) {
            // One partial per 16-lane row: lane r of the row ends with the row
            // total of value r = SEG_PMAX slot-major (slot, observable) pairs.
            double v[16];
#pragma unroll
            for (int k = 0; k < 16; k++) {
                v[k] = k < SEG_PMAX * NJ ? acc[k / NA][k % NA] : 0.0;
            }
            const double tot = row_reduce_scatter16(v);
            const int r = threadIdx.x & 15;
            const int s = r / NJ;
            if (r < SEG_PMAX * NJ && s < bd.nslots) {
                const int row = threadIdx.x >> 4;
                lacc[(row * pass_slots + bd.slot0 + s) * NJ + r % NJ] += tot;
            }
        }
    } else if constexpr (NJ > 0) {
        if (bd.nslots > 0) {
            // One partial per 16-lane row; each row's last lane owns its slot,
            // so the accumulation order is fixed.
            const bool owner = (threadIdx.x & 15) == 15;
            const int row = threadIdx.x >> 4;
#pragma unroll
            for (int s = 0; s < SEG_PMAX; s++) {
                if (s < bd.nslots) {
#pragma unroll
                    for (int j = 0; j < NJ; j++) {
                        const double v = row_total16(acc[s][j]);
                        if (owner) {
                            lacc[(row * pass_slots + bd.slot0 + s) * NJ + j] += v;
                        }
                    }
                }
            }
        }
    }
}

template <typename T, int NJ, typename Off>
__global__ void __launch_bounds__(TILE_BLOCK)
    tile_kernel(Cx<T> *__restrict__ phi, SegPtrs<T> hp, const TileBundle *__restrict__ bundles,
                const unsigned char *__restrict__ ops, TileLaunch L, double *__restrict__ partials) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NS = NJ + 1;
    const int TS = 1 << L.T;
    Cx<T> *tile = reinterpret_cast<Cx<T> *>(smem);
    double *lacc = reinterpret_cast<double *>(smem + static_cast<size_t>(NS) * TS * sizeof(Cx<T>));
    const int nacc = TILE_ROWS * L.nslots * (NJ > 0 ? NJ : 1);
    for (int i = threadIdx.x; i < nacc; i += TILE_BLOCK) {
        lacc[i] = 0.0;
    }
    Cx<T> *sp[NS];
    sp[0] = phi;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        sp[j + 1] = hp.p[j];
    }
    const uint64_t ntiles = uint64_t{1} << (L.n - L.T);
    // Tile position t = threadIdx.x + TILE_BLOCK * k: the global offset of
    // t's low bits is fixed per thread, that of k is uniform.
    uint64_t off_thread = 0;
    for (int q = 0; q < kBlockBits && q < L.T; q++) {
        off_thread |= static_cast<uint64_t>((threadIdx.x >> q) & 1) << L.tb[q];
    }
    const auto *cb = reinterpret_cast<const __attribute__((address_space(4))) TileBundle *>(
        (const __attribute__((address_space(4))) char *)(bundles));
    for (uint64_t g = blockIdx.x; g < ntiles; g += gridDim.x) {
        uint64_t base = g;
        for (int q = 0; q < L.T; q++) {
            base = insert_zero(base, L.tb[q]);
        }
        for (int t = threadIdx.x, k = 0; t < TS; t += TILE_BLOCK, k++) {
            const uint64_t off = off_thread | tile_high_offset(L, k);
            const Off bo = static_cast<Off>((base + off) * sizeof(Cx<T>));
#pragma unroll
            for (int s = 0; s < NS; s++) {
                tile[s * TS + t] = ldo(sp[s], bo);
            }
        }
        __syncthreads();
        for (int b = 0; b < L.nbundles; b++) {
            const TileBundle bd = {cb[b].K,     cb[b].qa,     cb[b].qb,    cb[b].first,
                                   cb[b].nops,  cb[b].nslots, cb[b].slot0, 0};
            if (bd.K == 1) {
                run_bundle<T, 1, NJ>(tile, TS, bd, ops, lacc, L.nslots);
            } else {
                run_bundle<T, 2, NJ>(tile, TS, bd, ops, lacc, L.nslots);
            }
            __syncthreads();
        }
        for (int t = threadIdx.x, k = 0; t < TS; t += TILE_BLOCK, k++) {
            const uint64_t off = off_thread | tile_high_offset(L, k);
            const Off bo = static_cast<Off>((base + off) * sizeof(Cx<T>));
            if (L.write_phi) {
                sto(sp[0], bo, tile[t]);
            }
#pragma unroll
            for (int s = 1; s < NS; s++) {
                sto(sp[s], bo, tile[s * TS + t]);
            }
        }
        __syncthreads();
    }
    if constexpr (NJ > 0) {
        const int per = L.nslots * NJ;
        for (int i = threadIdx.x; i < per; i += TILE_BLOCK) {
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < TILE_ROWS; w++) {
                v += lacc[w * per + i];
            }
            partials[static_cast<uint64_t>(i) * gridDim.x + blockIdx.x] = v;
        }
    }
}

constexpr int kReduceMax = TILE_PMAX > RT_PMAX ? TILE_PMAX : RT_PMAX;

struct TileReduceArgs {
    int grid;
    int nj;
    int j0;
    int ld;
    int col[kReduceMax];
    double coef[kReduceMax];
};

// One block per (slot, j): fixed-order FP64 sum of the per-block partials.
__global__ void __launch_bounds__(256)
    tile_reduce_kernel(const double *__restrict__ partials, TileReduceArgs a,
                       double *__restrict__ jac) {
    const int sj = blockIdx.x;
    const int s = sj / a.nj;
    const int j = sj % a.nj;
    double v = 0.0;
    for (int b = threadIdx.x; b < a.grid; b += blockDim.x) {
        v += partials[static_cast<uint64_t>(sj) * a.grid + b];
    }
    __shared__ double red[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        jac[static_cast<uint64_t>(a.j0 + j) * a.ld + a.col[s]] =
            a.coef[s] * (red[0] + red[1] + red[2] + red[3]);
    }
}

template <typename T, int NJ, typename Off>
void set_lds_limit() {
    static bool done = false;
    if (!done) {
        PLK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&tile_kernel<T, NJ, Off>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(TILE_LDS_MAX)));
        done = true;
    }
}

// Blocks of `kernel` that fit on one CU with `lds` bytes of dynamic LDS
// (registers and LDS both counted by the runtime); PLK_TILE_BLOCKS_PER_CU
// overrides. A grid of exactly the resident blocks avoids a second, partial
// round of blocks in the grid-stride loop.
int resident_blocks(const void *kernel, size_t lds) {
    const char *e = std::getenv("PLK_TILE_BLOCKS_PER_CU");
    if (e && *e) {
        return std::max(1, std::min(std::atoi(e), 8));
    }
    static std::map<std::pair<const void *, size_t>, int> cache;
    const auto key = std::make_pair(kernel, lds);
    auto it = cache.find(key);
    if (it != cache.end()) {
        return it->second;
    }
    int nb = 0;
    PLK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, TILE_BLOCK, lds));
    nb = std::max(1, std::min(nb, 8));
    cache[key] = nb;
    return nb;
}

template <typename T, int NJ>
int tile_launch_t(Device &dev, const State &phi, void *const *h_ptrs, const void *d_bundles,
                  const void *d_ops, const TileLaunch &L) {
    SegPtrs<T> hp{};
    for (int j = 0; j < NJ; j++) {
        hp.p[j] = static_cast<Cx<T> *>(h_ptrs[j]);
    }
    const size_t lds = tile_lds_bytes(phi.dtype, L.T, NJ, L.nslots);
    if (lds > TILE_LDS_MAX) {
        raise(PLK_ERR_INVALID, "tile pass exceeds the LDS budget");
    }
    const uint64_t ntiles = uint64_t{1} << (L.n - L.T);
    const bool small = phi.bytes() <= (uint64_t{1} << 32);
    const void *kfn = small ? reinterpret_cast<const void *>(&tile_kernel<T, NJ, uint32_t>)
                            : reinterpret_cast<const void *>(&tile_kernel<T, NJ, uint64_t>);
    if (small) {
        set_lds_limit<T, NJ, uint32_t>();
    } else {
        set_lds_limit<T, NJ, uint64_t>();
    }
    const uint64_t cap = static_cast<uint64_t>(dev.num_cus) * resident_blocks(kfn, lds);
    const unsigned grid = static_cast<unsigned>(std::min(ntiles, cap));
    if (NJ > 0 && L.nslots > 0) {
        ensure_red(dev, static_cast<size_t>(grid) * L.nslots * NJ, 0);
    }
    const double bytes = (1.0 + L.write_phi + 2.0 * NJ) * static_cast<double>(phi.bytes());
    ProfScope prof(NJ > 0 ? "adjoint_tile" : "apply_tile", bytes, dev.stream);
    if (phi.bytes() <= (uint64_t{1} << 32)) {
        set_lds_limit<T, NJ, uint32_t>();
        hipLaunchKernelGGL((tile_kernel<T, NJ, uint32_t>), dim3(grid), dim3(TILE_BLOCK), lds,
                           dev.stream, static_cast<Cx<T> *>(phi.d), hp,
                           static_cast<const TileBundle *>(d_bundles),
                           static_cast<const unsigned char *>(d_ops), L, dev.red);
    } else {
        set_lds_limit<T, NJ, uint64_t>();
        hipLaunchKernelGGL((tile_kernel<T, NJ, uint64_t>), dim3(grid), dim3(TILE_BLOCK), lds,
                           dev.stream, static_cast<Cx<T> *>(phi.d), hp,
                           static_cast<const TileBundle *>(d_bundles),
                           static_cast<const unsigned char *>(d_ops), L, dev.red);
    }
    PLK_HIP(hipGetLastError());
    return static_cast<int>(grid);
}

template <typename T>
int tile_launch_dtype(Device &dev, const State &phi, void *const *h, const void *b,
                      const void *o, const TileLaunch &L) {
    switch (L.nj) {
    case 0: return tile_launch_t<T, 0>(dev, phi, h, b, o, L);
    case 1: return tile_launch_t<T, 1>(dev, phi, h, b, o, L);
    case 2: return tile_launch_t<T, 2>(dev, phi, h, b, o, L);
    case 3: return tile_launch_t<T, 3>(dev, phi, h, b, o, L);
    case 4: return tile_launch_t<T, 4>(dev, phi, h, b, o, L);
    case 5:
        if constexpr (sizeof(T) == 4) {
            return tile_launch_t<T, 5>(dev, phi, h, b, o, L);
        }
        break;
    case 6:
        if constexpr (sizeof(T) == 4) {
            return tile_launch_t<T, 6>(dev, phi, h, b, o, L);
        }
        break;
    case 8:
        if constexpr (sizeof(T) == 4) {
            return tile_launch_t<T, 8>(dev, phi, h, b, o, L);
        }
        break;
    default:
        break;
    }
    raise(PLK_ERR_INVALID, "unsupported tile observable chunk " + std::to_string(L.nj));
}

} // namespace

size_t tile_lds_bytes(int dtype, int T, int nj, int nslots) {
    const size_t elem = dtype == PLK_C64 ? 8 : 16;
    return static_cast<size_t>(nj + 1) * (size_t{1} << T) * elem +
           static_cast<size_t>(TILE_ROWS) * nslots * (nj > 0 ? nj : 1) * sizeof(double);
}

bool tile_nj_supported(int dtype, int nj) {
    return (nj >= 0 && nj <= 4) || (dtype == PLK_C64 && (nj == 5 || nj == 6 || nj == 8));
}

int launch_tile(Device &dev, const State &phi, void *const *h_ptrs, const void *d_bundles,
                const void *d_ops, const TileLaunch &L) {
    if (L.T > TILE_TMAX || L.T > L.n || L.T < 1) {
        raise(PLK_ERR_INVALID, "tile width out of range");
    }
    return phi.dtype == PLK_C64 ? tile_launch_dtype<float>(dev, phi, h_ptrs, d_bundles, d_ops, L)
                                : tile_launch_dtype<double>(dev, phi, h_ptrs, d_bundles, d_ops, L);
}

void launch_tile_reduce(Device &dev, int grid, int nslots, int nj, int j0, const int *cols,
                        const double *coefs, double *d_jac, int ld) {
    if (nslots > kReduceMax) {
        raise(PLK_ERR_INVALID, "too many trainable operations in a tile pass");
    }
    TileReduceArgs a{};
    a.grid = grid;
    a.nj = nj;
    a.j0 = j0;
    a.ld = ld;
    for (int s = 0; s < nslots; s++) {
        a.col[s] = cols[s];
        a.coef[s] = coefs[s];
    }
    hipLaunchKernelGGL(tile_reduce_kernel, dim3(nslots * nj), dim3(256), 0, dev.stream, dev.red,
                       a, d_jac);
    PLK_HIP(hipGetLastError());
}

} // namespace plk
