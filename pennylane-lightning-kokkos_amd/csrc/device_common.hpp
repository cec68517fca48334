// Device-side building blocks shared by the gate, sweep and tile kernels:
// complex arithmetic, state element access, register-level application of
// one- and two-wire operations to a group of 2^K amplitudes, generator inner
// products and wave reductions.
#pragma once

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

// Value of lane (i - SHIFT) of the same 16-lane row, 0 when that lane is
// outside the row (DPP row_shr, VALU only).
template <int SHIFT> __device__ __forceinline__ double row_shr(double v) {
    static_assert(SHIFT >= 1 && SHIFT <= 15, "row_shr shift");
    const uint64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b & 0xffffffffu), 0x110 + SHIFT,
                                               0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), 0x110 + SHIFT, 0xF,
                                               0xF, true);
    return __longlong_as_double(static_cast<long long>(
        (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo)));
}

// Inclusive prefix sum within each 16-lane row: the row's last lane
// (lane % 16 == 15) receives the row total.
__device__ __forceinline__ double row_total16(double v) {
    v += row_shr<1>(v);
    v += row_shr<2>(v);
    v += row_shr<4>(v);
    v += row_shr<8>(v);
    return v;
}

// Value of lane CTRL-partner within the same 16-lane row for the DPP
// permutations row_mirror (0x140), row_half_mirror (0x141) and quad_perm
// (0x00-0xFF), whose sources are always inside the row.
template <int CTRL> __device__ __forceinline__ double dpp_partner(double v) {
    const uint64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b & 0xffffffffu), CTRL, 0xF,
                                               0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xF, 0xF,
                                               false);
    return __longlong_as_double(static_cast<long long>(
        (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo)));
}

template <int CTRL, int H>
__device__ __forceinline__ void butterfly_step(double (&v)[16], bool upper) {
#pragma unroll
    for (int k = 0; k < H; k++) {
        const double lo = v[k], hi = v[k + H];
        v[k] = (upper ? hi : lo) + dpp_partner<CTRL>(upper ? lo : hi);
    }
}

// Reduce-scatter of 16 values over each 16-lane row: afterwards lane r
// (lane % 16) holds the row total of value r. Partners lane^15, ^7, ^3, ^1
// (row_mirror, row_half_mirror, quad_perm [3,2,1,0], [1,0,3,2]); at each
// step a lane keeps the half selected by its bit 3, 2, 1, 0. The tree is
// fixed, so the sums are deterministic; 15 exchange-adds per lane instead of
// 4 per value for a 16-lane scan of each value.
__device__ __forceinline__ double row_reduce_scatter16(double (&v)[16]) {
    const int r = static_cast<int>(__lane_id()) & 15;
    butterfly_step<0x140, 8>(v, (r >> 3) & 1);
    butterfly_step<0x141, 4>(v, (r >> 2) & 1);
    butterfly_step<0x1B, 2>(v, (r >> 1) & 1);
    butterfly_step<0xB1, 1>(v, r & 1);
    return v[0];
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

template <typename T> struct SegPtrs {
    Cx<T> *p[SEG_NJMAX];
};

// Local index (within the op) of element i of a group: one-wire ops use bit
// Q, two-wire ops r = 2 * bit(QA) + bit(QB).
template <int M, int QA, int QB> __device__ __forceinline__ constexpr int op_index(int i) {
    return M == 1 ? ((i >> QA) & 1) : ((((i >> QA) & 1) << 1) | ((i >> QB) & 1));
}

// A register copy the compiler must keep where it is written. A permutation
// writes its out-of-place result as plain copies of its source; left to
// itself the compiler places those copies ahead of the branch on the gate
// structure, so every other structure's path executes them and then
// overwrites them. As instructions of their own they stay in the
// permutation's branch.
template <typename T> __device__ __forceinline__ Cx<T> kept_copy(Cx<T> x) {
    Cx<T> y;
    if constexpr (sizeof(T) == 8) {
        asm("v_mov_b64 %0, %1" : "=v"(y.r) : "v"(x.r));
        asm("v_mov_b64 %0, %1" : "=v"(y.i) : "v"(x.i));
    } else {
        asm("v_mov_b32 %0, %1" : "=v"(y.r) : "v"(x.r));
        asm("v_mov_b32 %0, %1" : "=v"(y.i) : "v"(x.i));
    }
    return y;
}

// dst = src with the elements whose op-local indices are A and B exchanged.
template <typename T, int K, int M, int QA, int QB, int A, int B>
__device__ __forceinline__ void permute_to(Cx<T> (&dst)[1 << K], const Cx<T> (&src)[1 << K]) {
#pragma unroll
    for (int i = 0; i < (1 << K); i++) {
        const int r = op_index<M, QA, QB>(i);
        int j = i;
        if (r == A || r == B) {
            if (M == 1) {
                j = i ^ (1 << QA);
            } else {
                if ((A ^ B) & 2) {
                    j ^= 1 << QA;
                }
                if ((A ^ B) & 1) {
                    j ^= 1 << QB;
                }
            }
        }
        dst[i] = kept_copy(src[j]);
    }
}

// dst = U src for one group, one helper per structure of U. src and dst are
// distinct arrays, so no result aliases an input still to be read.
template <typename T, int K, int M, int QA, int QB>
__device__ __forceinline__ void diag_to(Cx<T> (&dst)[1 << K], const Cx<T> (&src)[1 << K],
                                        const __attribute__((address_space(4))) T *U) {
#pragma unroll
    for (int i = 0; i < (1 << K); i++) {
        const int r = op_index<M, QA, QB>(i);
        dst[i] = cmul(Cx<T>{U[2 * r], U[2 * r + 1]}, src[i]);
    }
}

template <typename T, int K, int M, int QA, int QB>
__device__ __forceinline__ void real_to(Cx<T> (&dst)[1 << K], const Cx<T> (&src)[1 << K],
                                        const __attribute__((address_space(4))) T *U) {
    constexpr int D = 1 << K;
    if constexpr (M == 1) {
        const T m00 = U[0], m01 = U[1], m10 = U[2], m11 = U[3];
#pragma unroll
        for (int i = 0; i < D; i++) {
            if (i & (1 << QA)) {
                continue;
            }
            const int j = i | (1 << QA);
            const Cx<T> a = src[i], b = src[j];
            dst[i] = {fma(m00, a.r, m01 * b.r), fma(m00, a.i, m01 * b.i)};
            dst[j] = {fma(m10, a.r, m11 * b.r), fma(m10, a.i, m11 * b.i)};
        }
    } else {
#pragma unroll
        for (int i = 0; i < D; i++) {
            if (i & ((1 << QA) | (1 << QB))) {
                continue;
            }
            const int idx[4] = {i, i | (1 << QB), i | (1 << QA), i | (1 << QA) | (1 << QB)};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                T sr = U[4 * r] * src[idx[0]].r, si = U[4 * r] * src[idx[0]].i;
#pragma unroll
                for (int c = 1; c < 4; c++) {
                    sr = fma(U[4 * r + c], src[idx[c]].r, sr);
                    si = fma(U[4 * r + c], src[idx[c]].i, si);
                }
                dst[idx[r]] = {sr, si};
            }
        }
    }
}

template <typename T, int K, int M, int QA, int QB>
__device__ __forceinline__ void dense_to(Cx<T> (&dst)[1 << K], const Cx<T> (&src)[1 << K],
                                         const __attribute__((address_space(4))) T *U) {
    if constexpr (M == 1) {
        Cx<T> u[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            u[e] = {U[2 * e], U[2 * e + 1]};
        }
        gen1<T, K, QA>(dst, src, u);
    } else {
        Cx<T> u[16];
#pragma unroll
        for (int e = 0; e < 16; e++) {
            u[e] = {U[2 * e], U[2 * e + 1]};
        }
        gen2<T, K, QA, QB>(dst, src, u);
    }
}

// (fd, hd[j]) = U (fs, hs[j]) for phi and the NJ observable states. The
// structure of U is tested once per op and each branch applies it to every
// state, so the states' independent FP64 chains share one basic block and
// interleave in the schedule.
template <typename T, int K, int M, int QA, int QB, int NJ>
__device__ __forceinline__ void apply_kind_all(Cx<T> (&fd)[1 << K],
                                               Cx<T> (&hd)[NJ > 0 ? NJ : 1][1 << K],
                                               const Cx<T> (&fs)[1 << K],
                                               const Cx<T> (&hs)[NJ > 0 ? NJ : 1][1 << K],
                                               int ukind, int swap,
                                               const __attribute__((address_space(4))) T *U) {
    auto all = [&](auto &&fn) {
        fn(fd, fs);
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            fn(hd[j], hs[j]);
        }
    };
    using C = Cx<T>[1 << K];
    if (ukind == U_DIAG) {
        all([&](C &d, const C &x) { diag_to<T, K, M, QA, QB>(d, x, U); });
    } else if (ukind == U_SWAP) {
        if constexpr (M == 1) {
            all([&](C &d, const C &x) { permute_to<T, K, M, QA, QB, 0, 1>(d, x); });
        } else {
            switch (swap) {
            case 1: all([&](C &d, const C &x) { permute_to<T, K, M, QA, QB, 0, 1>(d, x); }); break;
            case 2: all([&](C &d, const C &x) { permute_to<T, K, M, QA, QB, 0, 2>(d, x); }); break;
            case 3: all([&](C &d, const C &x) { permute_to<T, K, M, QA, QB, 0, 3>(d, x); }); break;
            case 6: all([&](C &d, const C &x) { permute_to<T, K, M, QA, QB, 1, 2>(d, x); }); break;
            case 7: all([&](C &d, const C &x) { permute_to<T, K, M, QA, QB, 1, 3>(d, x); }); break;
            default: all([&](C &d, const C &x) { permute_to<T, K, M, QA, QB, 2, 3>(d, x); }); break;
            }
        }
    } else if (ukind == U_REAL) {
        all([&](C &d, const C &x) { real_to<T, K, M, QA, QB>(d, x, U); });
    } else {
        all([&](C &d, const C &x) { dense_to<T, K, M, QA, QB>(d, x, U); });
    }
}

// gf = G f for the group according to the generator's structure.
template <typename T, int K, int M, int QA, int QB>
__device__ __forceinline__ void gen_kind(Cx<T> (&gf)[1 << K], const Cx<T> (&f)[1 << K], int gkind,
                                         const __attribute__((address_space(4))) T *G) {
    constexpr int D = 1 << K;
    if (gkind == G_Z) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            const int r = op_index<M, QA, QB>(i);
            gf[i] = (__builtin_popcount(r) & 1) ? Cx<T>{-f[i].r, -f[i].i} : f[i];
        }
    } else if (gkind == G_DIAG) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            const int r = op_index<M, QA, QB>(i);
            const T g = G[r];
            gf[i] = {g * f[i].r, g * f[i].i};
        }
    } else if (M == 1 && gkind == G_X) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            gf[i] = f[i ^ (1 << QA)];
        }
    } else if (M == 1 && gkind == G_Y) {
        // Y = [[0, -i], [i, 0]]: (Y f)_0 = -i f_1, (Y f)_1 = i f_0.
#pragma unroll
        for (int i = 0; i < D; i++) {
            const Cx<T> p = f[i ^ (1 << QA)];
            gf[i] = ((i >> QA) & 1) ? Cx<T>{-p.i, p.r} : Cx<T>{p.i, -p.r};
        }
    } else {
        if constexpr (M == 1) {
            Cx<T> g[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                g[e] = {G[2 * e], G[2 * e + 1]};
            }
            gen1<T, K, QA>(gf, f, g);
        } else {
            Cx<T> g[16];
#pragma unroll
            for (int e = 0; e < 16; e++) {
                g[e] = {G[2 * e], G[2 * e + 1]};
            }
            gen2<T, K, QA, QB>(gf, f, g);
        }
    }
}

// One op of a group: (fs, hs) -> (fd, hd). For a trainable op, the FP64
// inner products Im<hs_j | G fs> enter the register queue q (newest first).
template <typename T, int K, int NJ>
__device__ __forceinline__ void op_step(const Cx<T> (&fs)[1 << K],
                                        const Cx<T> (&hs)[NJ > 0 ? NJ : 1][1 << K],
                                        Cx<T> (&fd)[1 << K], Cx<T> (&hd)[NJ > 0 ? NJ : 1][1 << K],
                                        double (&q)[SEG_PMAX][NJ > 0 ? NJ : 1], int &nt,
                                        const __attribute__((address_space(4))) unsigned char *base,
                                        int o) {
    constexpr int D = 1 << K;
    using CT = __attribute__((address_space(4))) const T;
    const auto *heads = reinterpret_cast<const __attribute__((address_space(4))) OpHead *>(base);
    const int cls = heads[o].cls;
    const int slot = heads[o].slot;
    const int kinds = heads[o].kinds;
    const uint32_t offs = static_cast<uint32_t>(heads[o].offs);
    const int ukind = kinds & 255;
    const int gkind = (kinds >> 8) & 255;
    const int swp = kinds >> 16;
    CT *U = reinterpret_cast<CT *>(base + (offs & 0xffff));
    CT *G = reinterpret_cast<CT *>(base + (offs >> 16));
    dispatch_cls<K>(cls, [&](auto Mc, auto QAc, auto QBc) {
        constexpr int m = decltype(Mc)::value;
        constexpr int qa = decltype(QAc)::value;
        constexpr int qb = decltype(QBc)::value;
        if constexpr (NJ > 0) {
            if (slot >= 0) {
                Cx<T> gf[D];
                gen_kind<T, K, m, qa, qb>(gf, fs, gkind, G);
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const double d = im_dot(hs[j], gf);
#pragma unroll
                    for (int s = SEG_PMAX - 1; s > 0; s--) {
                        q[s][j] = q[s - 1][j];
                    }
                    q[0][j] = d;
                }
                nt++;
            }
        }
        apply_kind_all<T, K, m, qa, qb, NJ>(fd, hd, fs, hs, ukind, swp, U);
    });
}

// Applies nops fused operations (descriptors in device memory, read through
// the constant address space so the uniform loads become scalar loads) to
// the register group f of phi and the groups h[j] of NJ observable states.
// Ops alternate between (f, h) and a second register set, so each result
// lands in registers whose old contents are dead and the loop carries no
// copies. For a trainable op the FP64 inner products Im<h_j | G f> are added
// to acc[slot][j] before the op is applied. The trainable ops of a group hold
// slots 0, 1, 2, ... in op order, so their products enter a register queue
// (newest first) and reach acc[slot] once per group, after the loop: the loop
// body needs no selects over the slots.
template <typename T, int K, int NJ>
__device__ __forceinline__ void run_ops(Cx<T> (&f)[1 << K], Cx<T> (&h)[NJ > 0 ? NJ : 1][1 << K],
                                        double (&acc)[SEG_PMAX][NJ > 0 ? NJ : 1],
                                        const void *group, int nops) {
    constexpr int D = 1 << K;
    constexpr int NA = NJ > 0 ? NJ : 1;
    const auto *base = (const __attribute__((address_space(4))) unsigned char *)(group);
    double q[SEG_PMAX][NA];
#pragma unroll
    for (int s = 0; s < SEG_PMAX; s++) {
#pragma unroll
        for (int j = 0; j < NA; j++) {
            q[s][j] = 0.0;
        }
    }
    int nt = 0;
    Cx<T> fb[D];
    Cx<T> hb[NA][D];
    for (int o = 0; o < nops; o += 2) {
        op_step<T, K, NJ>(f, h, fb, hb, q, nt, base, o);
        if (o + 1 < nops) {
            op_step<T, K, NJ>(fb, hb, f, h, q, nt, base, o + 1);
        } else {
#pragma unroll
            for (int i = 0; i < D; i++) {
                f[i] = fb[i];
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    h[j][i] = hb[j][i];
                }
            }
        }
    }
    if constexpr (NJ > 0) {
        // Align the queue so that q[SEG_PMAX - 1 - s] holds slot s.
#pragma unroll
        for (int r = 0; r < SEG_PMAX; r++) {
            if (r >= nt) {
#pragma unroll
                for (int j = 0; j < NJ; j++) {
#pragma unroll
                    for (int s = SEG_PMAX - 1; s > 0; s--) {
                        q[s][j] = q[s - 1][j];
                    }
                    q[0][j] = 0.0;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < SEG_PMAX; s++) {
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                acc[s][j] += q[SEG_PMAX - 1 - s][j];
            }
        }
    }
}

} // namespace plk
