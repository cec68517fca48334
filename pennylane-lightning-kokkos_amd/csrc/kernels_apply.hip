// Gate-application and adjoint-sweep kernels for gfx950 (CDNA4).
//
// The segment kernel is the hot path. A segment is a run of consecutive
// operations whose wires all lie in a set W of K <= 3 wires. Every thread
// owns one group of 2^K amplitudes (all index bits outside W fixed) of the
// state phi and of nj observable-applied states h_j, keeps them in registers
// while the whole run is applied, and accumulates the FP64 inner products
// Im<h_j| G phi> of the run's trainable operations. One launch is therefore
// one HBM pass: 2 * (1 + nj) * S bytes for a state of S bytes, however many
// operations the run holds. This replaces, per operation, the reference's
// mu copy, gate functor, generator functor and blocking inner-product reduce
// (algorithms/AdjointDiffKokkos.hpp:454-476; util/LinearAlgebraKokkos.hpp:
// 199-236).
#include <cstdio>

#include "device_common.hpp"

namespace plk {

template <typename T, int K, int NJ, typename Off>
__global__ void __launch_bounds__(SEG_BLOCK)
    seg_kernel(Cx<T> *__restrict__ phi, SegPtrs<T> hp,
               const unsigned char *__restrict__ ops, SegLaunch L, uint64_t ngroups,
               double *__restrict__ partials) {
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
    Off off[D];
#pragma unroll
    for (int i = 0; i < D; i++) {
        uint64_t o = 0;
#pragma unroll
        for (int q = 0; q < K; q++) {
            if (i & (1 << q)) {
                o |= uint64_t{1} << L.bits[q];
            }
        }
        off[i] = static_cast<Off>(o * sizeof(Cx<T>));
    }
    Cx<T> *hptr[NA];
#pragma unroll
    for (int j = 0; j < NA; j++) {
        hptr[j] = hp.p[j];
    }

    const uint64_t stride = uint64_t{gridDim.x} * blockDim.x;
    for (uint64_t g = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; g < ngroups;
         g += stride) {
        uint64_t bidx = g;
#pragma unroll
        for (int q = 0; q < K; q++) {
            bidx = insert_zero(bidx, L.bits[q]);
        }
        const Off base = static_cast<Off>(bidx * sizeof(Cx<T>));
        Cx<T> f[D];
        Cx<T> h[NA][D];
#pragma unroll
        for (int i = 0; i < D; i++) {
            f[i] = ldo(phi, static_cast<Off>(base + off[i]));
        }
        if constexpr (NJ > 0) {
#pragma unroll
            for (int j = 0; j < NJ; j++) {
#pragma unroll
                for (int i = 0; i < D; i++) {
                    h[j][i] = ldo(hptr[j], static_cast<Off>(base + off[i]));
                }
            }
        }
        run_ops<T, K, NJ>(f, h, acc, ops, L.nops);
        if (L.write_phi) {
#pragma unroll
            for (int i = 0; i < D; i++) {
                sto(phi, static_cast<Off>(base + off[i]), f[i]);
            }
        }
        if constexpr (NJ > 0) {
#pragma unroll
            for (int j = 0; j < NJ; j++) {
#pragma unroll
                for (int i = 0; i < D; i++) {
                    sto(hptr[j], static_cast<Off>(base + off[i]), h[j][i]);
                }
            }
        }
    }

    if constexpr (NJ > 0) {
        __shared__ double red[SEG_BLOCK / 64][SEG_PMAX * NA];
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
#pragma unroll
        for (int s = 0; s < SEG_PMAX; s++) {
            if (s < L.nslots) {
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const double v = wave_sum(acc[s][j]);
                    if (lane == 0) {
                        red[wave][s * NJ + j] = v;
                    }
                }
            }
        }
        __syncthreads();
        const int t = threadIdx.x;
        if (t < L.nslots * NJ) {
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < SEG_BLOCK / 64; w++) {
                v += red[w][t];
            }
            partials[uint64_t{static_cast<uint32_t>(t)} * gridDim.x + blockIdx.x] = v;
        }
    }
}

struct SegReduceArgs {
    int grid;
    int nj;
    int j0;
    int ld;
    int col[SEG_PMAX];
    double coef[SEG_PMAX];
};

// One block per (slot, j): fixed-order FP64 sum of the per-block partials.
__global__ void __launch_bounds__(256)
    seg_reduce_kernel(const double *__restrict__ partials, SegReduceArgs a,
                      double *__restrict__ jac) {
    const int sj = blockIdx.x;
    const int s = sj / a.nj;
    const int j = sj % a.nj;
    double v = 0.0;
    for (int b = threadIdx.x; b < a.grid; b += blockDim.x) {
        v += partials[uint64_t{static_cast<uint32_t>(sj)} * a.grid + b];
    }
    __shared__ double red[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double tot = red[0] + red[1] + red[2] + red[3];
        jac[static_cast<uint64_t>(a.j0 + j) * a.ld + a.col[s]] = a.coef[s] * tot;
    }
}

// ---------------------------------------------------------------------------
// Dense 2^m x 2^m operation through LDS (m <= 8): 256 >> m groups per block,
// one row per thread.
// ---------------------------------------------------------------------------
struct DenseArgs {
    int m;
    int bits[8];
    uint64_t off[256];
    uint64_t ngroups;
};

template <typename T>
__global__ void __launch_bounds__(256)
    dense_kernel(Cx<T> *__restrict__ sv, const Cx<T> *__restrict__ mat,
                 const DenseArgs *__restrict__ a) {
    __shared__ Cx<T> buf[256];
    const int m = a->m;
    const int dim = 1 << m;
    const int gpb = 256 >> m;
    const int gl = threadIdx.x >> m;
    const int r = threadIdx.x & (dim - 1);
    const uint64_t ng = a->ngroups;
    for (uint64_t g0 = uint64_t{blockIdx.x} * gpb; g0 < ng; g0 += uint64_t{gridDim.x} * gpb) {
        const uint64_t g = g0 + gl;
        const bool active = g < ng;
        uint64_t base = g;
        for (int q = 0; q < m; q++) {
            base = insert_zero(base, a->bits[q]);
        }
        const uint64_t idx = base + a->off[r];
        if (active) {
            buf[threadIdx.x] = ld(sv, idx);
        }
        __syncthreads();
        if (active) {
            const Cx<T> *row = mat + static_cast<uint64_t>(r) * dim;
            const Cx<T> *v = buf + (gl << m);
            Cx<T> s = {0, 0};
            for (int c = 0; c < dim; c++) {
                s = cmad(row[c], v[c], s);
            }
            st(sv, idx, s);
        }
        __syncthreads();
    }
}

// MultiRZ-style diagonal: factor `even` or `odd` by the parity of (i & mask).
template <typename T>
__global__ void __launch_bounds__(256)
    parity_kernel(Cx<T> *__restrict__ sv, uint64_t len, uint64_t mask, Cx<T> even,
                  Cx<T> odd) {
    for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < len;
         i += uint64_t{gridDim.x} * blockDim.x) {
        const Cx<T> f = (__popcll(i & mask) & 1) ? odd : even;
        st(sv, i, cmul(ld(sv, i), f));
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
    fill_kernel(Cx<T> *__restrict__ sv, uint64_t len, uint64_t one_at) {
    for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < len;
         i += uint64_t{gridDim.x} * blockDim.x) {
        st(sv, i, Cx<T>{static_cast<T>(i == one_at ? 1 : 0), static_cast<T>(0)});
    }
}

__global__ void __launch_bounds__(256) zero_doubles_kernel(double *__restrict__ p, uint64_t len) {
    for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < len;
         i += uint64_t{gridDim.x} * blockDim.x) {
        p[i] = 0.0;
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
    scatter_kernel(Cx<T> *__restrict__ sv, const uint64_t *__restrict__ idx,
                   const Cx<T> *__restrict__ vals, uint64_t count) {
    for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < count;
         i += uint64_t{gridDim.x} * blockDim.x) {
        st(sv, idx[i], ld(vals, i));
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
    axpy_kernel(Cx<T> *__restrict__ y, const Cx<T> *__restrict__ x, uint64_t len,
                double alpha) {
    const T a = static_cast<T>(alpha);
    for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < len;
         i += uint64_t{gridDim.x} * blockDim.x) {
        const Cx<T> xv = ld(x, i);
        Cx<T> yv = ld(y, i);
        yv.r = fma(a, xv.r, yv.r);
        yv.i = fma(a, xv.i, yv.i);
        st(y, i, yv);
    }
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
namespace {

unsigned grid_for(const Device &dev, uint64_t work, unsigned per_block = 256) {
    const uint64_t want = (work + per_block - 1) / per_block;
    const uint64_t cap = static_cast<uint64_t>(dev.num_cus) * 8;
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min(want, cap)));
}

template <typename T, int K, int NJ>
int seg_launch_t(Device &dev, const State &phi, void *const *h_ptrs, const void *d_ops,
                 const SegLaunch &L) {
    SegPtrs<T> hp{};
    for (int j = 0; j < NJ; j++) {
        hp.p[j] = static_cast<Cx<T> *>(h_ptrs[j]);
    }
    const uint64_t ng = uint64_t{1} << (L.n - K);
    const unsigned grid = grid_for(dev, ng, SEG_BLOCK);
    if (NJ > 0 && L.nslots > 0) {
        ensure_red(dev, static_cast<size_t>(grid) * SEG_PMAX * SEG_NJMAX, 0);
    }
    // phi is read (and written when write_phi), every h_j is read and written.
    const double bytes = (1.0 + L.write_phi + 2.0 * NJ) * static_cast<double>(phi.bytes());
    ProfScope prof(NJ > 0 ? "adjoint_segment" : "apply_segment", bytes, dev.stream);
    if (phi.bytes() <= (uint64_t{1} << 32)) {
        hipLaunchKernelGGL((seg_kernel<T, K, NJ, uint32_t>), dim3(grid), dim3(SEG_BLOCK), 0,
                           dev.stream, static_cast<Cx<T> *>(phi.d), hp,
                           static_cast<const unsigned char *>(d_ops), L, ng, dev.red);
    } else {
        hipLaunchKernelGGL((seg_kernel<T, K, NJ, uint64_t>), dim3(grid), dim3(SEG_BLOCK), 0,
                           dev.stream, static_cast<Cx<T> *>(phi.d), hp,
                           static_cast<const unsigned char *>(d_ops), L, ng, dev.red);
    }
    PLK_HIP(hipGetLastError());
    return static_cast<int>(grid);
}

} // namespace

// Largest observable chunk per launch that stays in registers (no spills).
int seg_max_nj(int dtype, int K) {
    if (dtype == PLK_C64) {
        return K == 1 ? 8 : 4;
    }
    return K <= 2 ? 4 : 2;
}

namespace {

template <typename T, int K>
int seg_launch_k(Device &dev, const State &phi, void *const *h, const void *d_ops,
                 const SegLaunch &L) {
    switch (L.nj) {
    case 0: return seg_launch_t<T, K, 0>(dev, phi, h, d_ops, L);
    case 1: return seg_launch_t<T, K, 1>(dev, phi, h, d_ops, L);
    case 2: return seg_launch_t<T, K, 2>(dev, phi, h, d_ops, L);
    case 4:
        if constexpr (sizeof(T) == 4 || K < 3) {
            return seg_launch_t<T, K, 4>(dev, phi, h, d_ops, L);
        }
        break;
    case 8:
        if constexpr (K == 1 && sizeof(T) == 4) {
            return seg_launch_t<T, K, 8>(dev, phi, h, d_ops, L);
        }
        break;
    default:
        break;
    }
    raise(PLK_ERR_INVALID, "unsupported observable chunk " + std::to_string(L.nj));
}

template <typename T>
int seg_launch_dtype(Device &dev, const State &phi, void *const *h, const void *d_ops,
                     const SegLaunch &L) {
    switch (L.K) {
    case 1: return seg_launch_k<T, 1>(dev, phi, h, d_ops, L);
    case 2: return seg_launch_k<T, 2>(dev, phi, h, d_ops, L);
    case 3: return seg_launch_k<T, 3>(dev, phi, h, d_ops, L);
    default: raise(PLK_ERR_INVALID, "unsupported segment width");
    }
}

} // namespace

int launch_segment(Device &dev, const State &phi, void *const *h_ptrs, const void *d_ops,
                   const SegLaunch &L) {
    if (L.n < L.K) {
        raise(PLK_ERR_INVALID, "segment wider than the state");
    }
    return phi.dtype == PLK_C64 ? seg_launch_dtype<float>(dev, phi, h_ptrs, d_ops, L)
                                : seg_launch_dtype<double>(dev, phi, h_ptrs, d_ops, L);
}

void launch_segment_reduce(Device &dev, int grid, int nslots, int nj, int j0, const int *cols,
                           const double *coefs, double *d_jac, int ld) {
    SegReduceArgs a{};
    a.grid = grid;
    a.nj = nj;
    a.j0 = j0;
    a.ld = ld;
    for (int s = 0; s < nslots; s++) {
        a.col[s] = cols[s];
        a.coef[s] = coefs[s];
    }
    hipLaunchKernelGGL(seg_reduce_kernel, dim3(nslots * nj), dim3(256), 0, dev.stream, dev.red,
                       a, d_jac);
    PLK_HIP(hipGetLastError());
}

namespace {

template <typename T>
void dense_t(Device &dev, const State &s, const std::vector<size_t> &wires,
             const std::vector<cd> &mat) {
    const int m = static_cast<int>(wires.size());
    if (m > 8) {
        raise(PLK_ERR_UNSUPPORTED, "dense operations on more than 8 wires are not supported");
    }
    const size_t dim = size_t{1} << m;
    DenseArgs a{};
    a.m = m;
    std::vector<int> bits;
    for (auto w : wires) {
        bits.push_back(s.n - 1 - static_cast<int>(w));
    }
    std::vector<int> sorted = bits;
    std::sort(sorted.begin(), sorted.end());
    for (int q = 0; q < m; q++) {
        a.bits[q] = sorted[q];
    }
    for (size_t r = 0; r < dim; r++) {
        uint64_t o = 0;
        for (int p = 0; p < m; p++) {
            if ((r >> (m - 1 - p)) & 1U) {
                o |= uint64_t{1} << bits[p];
            }
        }
        a.off[r] = o;
    }
    a.ngroups = uint64_t{1} << (s.n - m);
    std::vector<Cx<T>> hm(dim * dim);
    for (size_t i = 0; i < dim * dim; i++) {
        hm[i] = {static_cast<T>(mat[i].real()), static_cast<T>(mat[i].imag())};
    }
    // Matrix and index table travel in one device block (stream ordered).
    const size_t mat_bytes = hm.size() * sizeof(Cx<T>);
    const size_t args_off = (mat_bytes + 255) / 256 * 256;
    const size_t total = args_off + sizeof(DenseArgs);
    std::vector<unsigned char> blob(total);
    std::memcpy(blob.data(), hm.data(), mat_bytes);
    std::memcpy(blob.data() + args_off, &a, sizeof(DenseArgs));
    void *d_blob = nullptr;
    PLK_HIP(hipMallocAsync(&d_blob, total, dev.stream));
    PLK_HIP(hipMemcpyAsync(d_blob, blob.data(), total, hipMemcpyHostToDevice, dev.stream));
    const uint64_t gpb = 256 >> m;
    const unsigned grid = grid_for(dev, a.ngroups, static_cast<unsigned>(gpb));
    {
        ProfScope prof("apply_dense", 2.0 * static_cast<double>(s.bytes()), dev.stream);
        hipLaunchKernelGGL((dense_kernel<T>), dim3(grid), dim3(256), 0, dev.stream,
                           static_cast<Cx<T> *>(s.d), reinterpret_cast<const Cx<T> *>(d_blob),
                           reinterpret_cast<const DenseArgs *>(static_cast<unsigned char *>(d_blob) +
                                                               args_off));
        PLK_HIP(hipGetLastError());
    }
    PLK_HIP(hipFreeAsync(d_blob, dev.stream));
    // The host blob must stay valid until the copy has consumed it.
    PLK_HIP(hipStreamSynchronize(dev.stream));
}

template <typename T>
void parity_t(Device &dev, const State &s, uint64_t mask, cd even, cd odd) {
    const unsigned grid = grid_for(dev, s.len);
    ProfScope prof("apply_parity", 2.0 * static_cast<double>(s.bytes()), dev.stream);
    hipLaunchKernelGGL((parity_kernel<T>), dim3(grid), dim3(256), 0, dev.stream,
                       static_cast<Cx<T> *>(s.d), static_cast<uint64_t>(s.len), mask,
                       Cx<T>{static_cast<T>(even.real()), static_cast<T>(even.imag())},
                       Cx<T>{static_cast<T>(odd.real()), static_cast<T>(odd.imag())});
    PLK_HIP(hipGetLastError());
}

} // namespace

void launch_dense_matrix(Device &dev, const State &s, const std::vector<size_t> &wires,
                         const std::vector<cd> &mat) {
    if (wires.size() > 8) {
        // Up to 12 wires: out of place into a temporary, then back.
        TmpState tmp(s.n, s.dtype, s.device);
        dense_apply_into(dev, tmp.s, s, wires, mat);
        launch_copy(dev, s, tmp.s);
        PLK_HIP(hipStreamSynchronize(dev.stream));
        return;
    }
    if (wires.size() >= 4 && wires.size() <= 5 && s.n >= 9) {
        // 16x16 and 32x32 blocks run on the matrix cores.
        CompiledOp op;
        op.name = "matrix";
        op.kind = OpKind::Dense;
        op.wires = wires;
        op.mat = mat;
        apply_blocks(dev, s, fuse_blocks({op}, s.n));
        return;
    }
    if (s.dtype == PLK_C64) {
        dense_t<float>(dev, s, wires, mat);
    } else {
        dense_t<double>(dev, s, wires, mat);
    }
}

void launch_dense(Device &dev, const State &s, const CompiledOp &op) {
    launch_dense_matrix(dev, s, op.wires, op.mat);
}

void launch_parity(Device &dev, const State &s, const std::vector<size_t> &wires, cd even,
                   cd odd) {
    uint64_t mask = 0;
    for (auto w : wires) {
        mask |= uint64_t{1} << (s.n - 1 - static_cast<int>(w));
    }
    if (s.dtype == PLK_C64) {
        parity_t<float>(dev, s, mask, even, odd);
    } else {
        parity_t<double>(dev, s, mask, even, odd);
    }
}

void launch_fill_zero(Device &dev, const State &s) {
    launch_set_basis(dev, s, ~uint64_t{0});
}

void launch_fill_zero_doubles(Device &dev, double *p, size_t len) {
    if (len == 0) {
        return;
    }
    const unsigned grid = grid_for(dev, len);
    hipLaunchKernelGGL(zero_doubles_kernel, dim3(grid), dim3(256), 0, dev.stream, p,
                       static_cast<uint64_t>(len));
    PLK_HIP(hipGetLastError());
}

void launch_set_basis(Device &dev, const State &s, size_t index) {
    const unsigned grid = grid_for(dev, s.len);
    if (s.dtype == PLK_C64) {
        hipLaunchKernelGGL((fill_kernel<float>), dim3(grid), dim3(256), 0, dev.stream,
                           static_cast<Cx<float> *>(s.d), static_cast<uint64_t>(s.len),
                           static_cast<uint64_t>(index));
    } else {
        hipLaunchKernelGGL((fill_kernel<double>), dim3(grid), dim3(256), 0, dev.stream,
                           static_cast<Cx<double> *>(s.d), static_cast<uint64_t>(s.len),
                           static_cast<uint64_t>(index));
    }
    PLK_HIP(hipGetLastError());
}

void launch_scatter(Device &dev, const State &s, const size_t *d_idx, const void *d_vals,
                    size_t count) {
    if (count == 0) {
        return;
    }
    const unsigned grid = grid_for(dev, count);
    if (s.dtype == PLK_C64) {
        hipLaunchKernelGGL((scatter_kernel<float>), dim3(grid), dim3(256), 0, dev.stream,
                           static_cast<Cx<float> *>(s.d), reinterpret_cast<const uint64_t *>(d_idx),
                           static_cast<const Cx<float> *>(d_vals), static_cast<uint64_t>(count));
    } else {
        hipLaunchKernelGGL((scatter_kernel<double>), dim3(grid), dim3(256), 0, dev.stream,
                           static_cast<Cx<double> *>(s.d), reinterpret_cast<const uint64_t *>(d_idx),
                           static_cast<const Cx<double> *>(d_vals), static_cast<uint64_t>(count));
    }
    PLK_HIP(hipGetLastError());
}

void launch_copy(Device &dev, const State &dst, const State &src) {
    if (dst.len != src.len || dst.dtype != src.dtype) {
        raise(PLK_ERR_INVALID, "state copy between different shapes");
    }
    ProfScope prof("copy", 2.0 * static_cast<double>(src.bytes()), dev.stream);
    PLK_HIP(hipMemcpyAsync(dst.d, src.d, src.bytes(), hipMemcpyDeviceToDevice, dev.stream));
}

void launch_axpy(Device &dev, const State &y, double alpha, const State &x) {
    const unsigned grid = grid_for(dev, y.len);
    if (y.dtype == PLK_C64) {
        hipLaunchKernelGGL((axpy_kernel<float>), dim3(grid), dim3(256), 0, dev.stream,
                           static_cast<Cx<float> *>(y.d), static_cast<const Cx<float> *>(x.d),
                           static_cast<uint64_t>(y.len), alpha);
    } else {
        hipLaunchKernelGGL((axpy_kernel<double>), dim3(grid), dim3(256), 0, dev.stream,
                           static_cast<Cx<double> *>(y.d), static_cast<const Cx<double> *>(x.d),
                           static_cast<uint64_t>(y.len), alpha);
    }
    PLK_HIP(hipGetLastError());
}

} // namespace plk
