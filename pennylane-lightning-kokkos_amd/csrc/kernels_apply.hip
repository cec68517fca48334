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

#include "plk_internal.hpp"

namespace plk {

template <typename T> struct Cx {
    T r, i;
};

template <typename T>
__device__ __forceinline__ Cx<T> cmul(const Cx<T> &a, const Cx<T> &b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}

// c + a * b
template <typename T>
__device__ __forceinline__ Cx<T> cmad(const Cx<T> &a, const Cx<T> &b,
                                      const Cx<T> &c) {
    return {fma(a.r, b.r, fma(-a.i, b.i, c.r)), fma(a.r, b.i, fma(a.i, b.r, c.i))};
}

// Im(conj(h) * g) in FP64.
template <typename T>
__device__ __forceinline__ double im_conj_mul(const Cx<T> &h, const Cx<T> &g) {
    return static_cast<double>(h.r) * static_cast<double>(g.i) -
           static_cast<double>(h.i) * static_cast<double>(g.r);
}

template <typename T> struct VecOf;
template <> struct VecOf<float> {
    using type = float2;
};
template <> struct VecOf<double> {
    using type = double2;
};

template <typename T>
__device__ __forceinline__ Cx<T> ld(const Cx<T> *p, uint64_t i) {
    using V = typename VecOf<T>::type;
    const V v = reinterpret_cast<const V *>(p)[i];
    return {v.x, v.y};
}

template <typename T>
__device__ __forceinline__ void st(Cx<T> *p, uint64_t i, const Cx<T> &c) {
    using V = typename VecOf<T>::type;
    V v;
    v.x = c.r;
    v.y = c.i;
    reinterpret_cast<V *>(p)[i] = v;
}

// Element access through a byte offset from a uniform base. With 32-bit
// offsets the compiler emits SGPR-base + VGPR-offset addressing, so one VGPR
// per amplitude position serves every state of the group.
template <typename T, typename Off>
__device__ __forceinline__ Cx<T> ldo(const Cx<T> *p, Off byte_off) {
    using V = typename VecOf<T>::type;
    const V v = *reinterpret_cast<const V *>(reinterpret_cast<const char *>(p) + byte_off);
    return {v.x, v.y};
}

template <typename T, typename Off>
__device__ __forceinline__ void sto(Cx<T> *p, Off byte_off, const Cx<T> &c) {
    using V = typename VecOf<T>::type;
    V v;
    v.x = c.r;
    v.y = c.i;
    *reinterpret_cast<V *>(reinterpret_cast<char *>(p) + byte_off) = v;
}

__device__ __forceinline__ uint64_t insert_zero(uint64_t k, int b) {
    const uint64_t low = k & ((uint64_t{1} << b) - 1);
    return ((k >> b) << (b + 1)) | low;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, 64);
    }
    return v;
}

// ---------------------------------------------------------------------------
// Register-level gate application on a group of 2^K amplitudes.
// One-wire op on local bit Q with 2x2 matrix u.
// ---------------------------------------------------------------------------
template <typename T, int K, int Q>
__device__ __forceinline__ void ap1(Cx<T> (&v)[1 << K], const Cx<T> (&u)[4]) {
#pragma unroll
    for (int i = 0; i < (1 << K); i++) {
        if (i & (1 << Q)) {
            continue;
        }
        const int j = i | (1 << Q);
        const Cx<T> a = v[i], b = v[j];
        v[i] = cmad(u[1], b, cmul(u[0], a));
        v[j] = cmad(u[3], b, cmul(u[2], a));
    }
}

// Two-wire op on local bits QA > QB; matrix index r = 2*bit(QA) + bit(QB).
template <typename T, int K, int QA, int QB>
__device__ __forceinline__ void ap2(Cx<T> (&v)[1 << K], const Cx<T> (&u)[16]) {
#pragma unroll
    for (int i = 0; i < (1 << K); i++) {
        if (i & ((1 << QA) | (1 << QB))) {
            continue;
        }
        const int idx[4] = {i, i | (1 << QB), i | (1 << QA), i | (1 << QA) | (1 << QB)};
        Cx<T> a[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            a[r] = v[idx[r]];
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            Cx<T> s = cmul(u[4 * r], a[0]);
#pragma unroll
            for (int c = 1; c < 4; c++) {
                s = cmad(u[4 * r + c], a[c], s);
            }
            v[idx[r]] = s;
        }
    }
}

// gf = G f for a one-wire generator on local bit Q (computed once per op).
template <typename T, int K, int Q>
__device__ __forceinline__ void gen1(Cx<T> (&gf)[1 << K], const Cx<T> (&f)[1 << K],
                                     const Cx<T> (&g)[4]) {
#pragma unroll
    for (int i = 0; i < (1 << K); i++) {
        if (i & (1 << Q)) {
            continue;
        }
        const int j = i | (1 << Q);
        gf[i] = cmad(g[1], f[j], cmul(g[0], f[i]));
        gf[j] = cmad(g[3], f[j], cmul(g[2], f[i]));
    }
}

template <typename T, int K, int QA, int QB>
__device__ __forceinline__ void gen2(Cx<T> (&gf)[1 << K], const Cx<T> (&f)[1 << K],
                                     const Cx<T> (&g)[16]) {
#pragma unroll
    for (int i = 0; i < (1 << K); i++) {
        if (i & ((1 << QA) | (1 << QB))) {
            continue;
        }
        const int idx[4] = {i, i | (1 << QB), i | (1 << QA), i | (1 << QA) | (1 << QB)};
#pragma unroll
        for (int r = 0; r < 4; r++) {
            Cx<T> t = cmul(g[4 * r], f[idx[0]]);
#pragma unroll
            for (int c = 1; c < 4; c++) {
                t = cmad(g[4 * r + c], f[idx[c]], t);
            }
            gf[idx[r]] = t;
        }
    }
}

// sum over the group of Im(conj(h) * gf), accumulated in FP64.
template <typename T, int D>
__device__ __forceinline__ double im_dot(const Cx<T> (&h)[D], const Cx<T> (&gf)[D]) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++) {
        s += im_conj_mul(h[i], gf[i]);
    }
    return s;
}

template <int V> using IC = std::integral_constant<int, V>;

// Maps the host-computed class of an op to (wire count, QA, QB).
// K=1: 0 -> (1,0). K=2: 0,1 -> one wire on bit 0/1; 2 -> (1,0).
// K=3: 0,1,2 -> one wire; 3 -> (1,0); 4 -> (2,0); 5 -> (2,1).
template <int K, typename F>
__device__ __forceinline__ void dispatch_cls(int cls, F &&f) {
    if constexpr (K == 1) {
        f(IC<1>{}, IC<0>{}, IC<0>{});
    } else if constexpr (K == 2) {
        switch (cls) {
        case 0: f(IC<1>{}, IC<0>{}, IC<0>{}); break;
        case 1: f(IC<1>{}, IC<1>{}, IC<0>{}); break;
        default: f(IC<2>{}, IC<1>{}, IC<0>{}); break;
        }
    } else {
        switch (cls) {
        case 0: f(IC<1>{}, IC<0>{}, IC<0>{}); break;
        case 1: f(IC<1>{}, IC<1>{}, IC<0>{}); break;
        case 2: f(IC<1>{}, IC<2>{}, IC<0>{}); break;
        case 3: f(IC<2>{}, IC<1>{}, IC<0>{}); break;
        case 4: f(IC<2>{}, IC<2>{}, IC<0>{}); break;
        default: f(IC<2>{}, IC<2>{}, IC<1>{}); break;
        }
    }
}

template <typename T>
using ConstSegOp = __attribute__((address_space(4))) const SegOp<T>;

template <typename T> struct SegPtrs {
    Cx<T> *p[SEG_NJMAX];
};

template <typename T, int K, int NJ, typename Off>
__global__ void __launch_bounds__(SEG_BLOCK)
    seg_kernel(Cx<T> *__restrict__ phi, SegPtrs<T> hp,
               const SegOp<T> *__restrict__ ops, SegLaunch L, uint64_t ngroups,
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
        for (int o = 0; o < L.nops; o++) {
            // Descriptors are read through the constant address space so that
            // the uniform loads become scalar loads into SGPRs.
            const ConstSegOp<T> &op = reinterpret_cast<const ConstSegOp<T> *>(
                (const __attribute__((address_space(4))) char *)(ops))[o];
            const int cls = op.cls;
            const int slot = op.slot;
            dispatch_cls<K>(cls, [&](auto M, auto QA, auto QB) {
                constexpr int m = decltype(M)::value;
                constexpr int qa = decltype(QA)::value;
                constexpr int qb = decltype(QB)::value;
                if constexpr (m == 1) {
                    if constexpr (NJ > 0) {
                        if (slot >= 0) {
                            Cx<T> gm[4];
#pragma unroll
                            for (int e = 0; e < 4; e++) {
                                gm[e] = {op.G[2 * e], op.G[2 * e + 1]};
                            }
                            Cx<T> gf[D];
                            gen1<T, K, qa>(gf, f, gm);
#pragma unroll
                            for (int j = 0; j < NJ; j++) {
                                const double d = im_dot(h[j], gf);
#pragma unroll
                                for (int s = 0; s < SEG_PMAX; s++) {
                                    if (s == slot) {
                                        acc[s][j] += d;
                                    }
                                }
                            }
                        }
                    }
                    Cx<T> u[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        u[e] = {op.U[2 * e], op.U[2 * e + 1]};
                    }
                    ap1<T, K, qa>(f, u);
                    if constexpr (NJ > 0) {
#pragma unroll
                        for (int j = 0; j < NJ; j++) {
                            ap1<T, K, qa>(h[j], u);
                        }
                    }
                } else {
                    if constexpr (NJ > 0) {
                        if (slot >= 0) {
                            Cx<T> gm[16];
#pragma unroll
                            for (int e = 0; e < 16; e++) {
                                gm[e] = {op.G[2 * e], op.G[2 * e + 1]};
                            }
                            Cx<T> gf[D];
                            gen2<T, K, qa, qb>(gf, f, gm);
#pragma unroll
                            for (int j = 0; j < NJ; j++) {
                                const double d = im_dot(h[j], gf);
#pragma unroll
                                for (int s = 0; s < SEG_PMAX; s++) {
                                    if (s == slot) {
                                        acc[s][j] += d;
                                    }
                                }
                            }
                        }
                    }
                    Cx<T> u[16];
#pragma unroll
                    for (int e = 0; e < 16; e++) {
                        u[e] = {op.U[2 * e], op.U[2 * e + 1]};
                    }
                    ap2<T, K, qa, qb>(f, u);
                    if constexpr (NJ > 0) {
#pragma unroll
                        for (int j = 0; j < NJ; j++) {
                            ap2<T, K, qa, qb>(h[j], u);
                        }
                    }
                }
            });
        }
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
                           static_cast<const SegOp<T> *>(d_ops), L, ng, dev.red);
    } else {
        hipLaunchKernelGGL((seg_kernel<T, K, NJ, uint64_t>), dim3(grid), dim3(SEG_BLOCK), 0,
                           dev.stream, static_cast<Cx<T> *>(phi.d), hp,
                           static_cast<const SegOp<T> *>(d_ops), L, ng, dev.red);
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
    return K == 1 ? 8 : (K == 2 ? 4 : 2);
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
        if constexpr (K == 1) {
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

// Explicit instantiation of the segment operand layout for the host side.
template struct SegOp<float>;
template struct SegOp<double>;

} // namespace plk
