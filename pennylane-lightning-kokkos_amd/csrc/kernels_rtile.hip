// Register-tile passes: many fused operations per HBM pass, with the tile in
// VGPRs instead of LDS.
//
// Each wavefront owns one tile at a time: the 2^(R+6) amplitudes whose index
// bits outside the pass's tile bits are fixed. Every lane holds 2^R of them
// per state (R register bits: 3 for complex128, 4 for complex64, so a lane
// holds 128 bytes of each state); the 6 lane bits select the lane. Which tile
// bit sits on which register bit or lane position is the pass's layout,
// chosen on the host and changed by remaps.
//
// A pass program is a list of 32-byte instructions read through the scalar
// cache. Each instruction first applies up to two remaps -- exchanges of a
// register bit with lane position 4 or 5, one v_permlane16/32_swap per
// 32-bit word pair, in place -- then one operation:
//   RT_REG1   one wire on register bit a;
//   RT_REG2   two wires on register bits a > b;
//   RT_SEL    mixing register bit a, with the op's other wire on lane position
//             s0: each lane applies the 2x2 block its lane bit selects
//             (controlled gates with the control on a lane, diagonal gates
//             across a register and a lane bit);
//   RT_LANE   diagonal on lane positions only: one factor per lane.
// A selector s0 / s1 of 64 + b names a global index bit b outside the tile
// instead of a lane: its value is fixed for the whole tile, so an operation
// whose non-mixing wires (a control, a diagonal factor) lie outside the tile
// needs no tile bit for them.
//
// Operations write out of place. The NS states (phi and NS - 1 observable
// states) live in NS + 1 register buffers: each operation writes state s into
// the buffer that is free at that moment, which leaves the buffer state s
// occupied free for state s + 1. After one instruction every state has moved
// one buffer down (state s in buffer (s - rot) mod (NS + 1) after rot
// instructions), so the instruction loop is unrolled NS + 1 times with the
// buffer indices as constants: no register copies between instructions and
// one spare buffer instead of a second copy of every state.
//
// A trainable operation forms G phi in the free buffer and adds the FP64
// products Im<h_j | G phi> of the lane's amplitudes; the four lanes of each
// quad are summed by DPP and one lane adds the total into LDS (ds_add_f64,
// one address per quad, so the order is fixed). After its last tile the
// wavefront writes per-block partials; a second kernel sums them in a fixed
// order. No LDS staging of states and no barriers: each wavefront streams its
// tiles independently, which keeps HBM busy while other wavefronts compute.
#include <map>
#include <utility>

#include "device_common.hpp"

namespace plk {

namespace {

template <typename T> using CT = const __attribute__((address_space(4))) T;

template <typename T> constexpr int rt_r() { return sizeof(T) == 8 ? 3 : 4; }

__device__ __forceinline__ bool lbit(int lane, int p) { return (lane >> p) & 1; }

// Block selector of RT_SEL / RT_LANE: a lane position (0-5), or 64 + a
// global index bit outside the tile, whose value is the same for the whole
// tile (tbase: the tile's first amplitude index).
__device__ __forceinline__ bool selbit(int lane, uint64_t tbase, int s) {
    return s < 64 ? lbit(lane, s) : ((tbase >> (s - 64)) & 1);
}

// Buffer of state s before (in) and after (out) instruction number `ROT`
// (mod NS + 1) of the rotation.
template <int NS, int ROT> struct Rot {
    static constexpr int NB = NS + 1;
    static constexpr int in(int s) { return ((s - ROT) % NB + NB) % NB; }
    static constexpr int out(int s) { return s == 0 ? ((NS - ROT) % NB + NB) % NB : in(s - 1); }
};

// ---------------------------------------------------------------------------
// Remap: exchange register bit Q with lane position P (4 or 5), in place.
// Afterwards lanes whose lane bit P is 1 hold the partner's b in a, and lanes
// whose bit is 0 hold the partner's a in b (v_permlane16_swap swaps the odd
// 16-lane rows of its first operand with the even rows of its second,
// v_permlane32_swap the upper half of the first with the lower half of the
// second).
// ---------------------------------------------------------------------------
template <int P> __device__ __forceinline__ void xchg_word(uint32_t &a, uint32_t &b) {
    static_assert(P == 4 || P == 5, "remaps use lane positions 4 and 5");
    if constexpr (P == 5) {
        const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    }
}

template <int P> __device__ __forceinline__ void xchg(double &a, double &b) {
    uint64_t ua = __double_as_longlong(a), ub = __double_as_longlong(b);
    uint32_t a0 = static_cast<uint32_t>(ua), a1 = static_cast<uint32_t>(ua >> 32);
    uint32_t b0 = static_cast<uint32_t>(ub), b1 = static_cast<uint32_t>(ub >> 32);
    xchg_word<P>(a0, b0);
    xchg_word<P>(a1, b1);
    a = __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(a1) << 32) | a0));
    b = __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(b1) << 32) | b0));
}

template <int P> __device__ __forceinline__ void xchg(float &a, float &b) {
    uint32_t ua = __float_as_uint(a), ub = __float_as_uint(b);
    xchg_word<P>(ua, ub);
    a = __uint_as_float(ua);
    b = __uint_as_float(ub);
}

template <typename T, int A, int Q, int P> __device__ __forceinline__ void remap1(Cx<T> (&v)[A]) {
#pragma unroll
    for (int k = 0; k < A; k++) {
        if (k & (1 << Q)) {
            continue;
        }
        xchg<P>(v[k].r, v[k | (1 << Q)].r);
        xchg<P>(v[k].i, v[k | (1 << Q)].i);
    }
}

// ---------------------------------------------------------------------------
// Per-lane selection among uniform blocks: block c = bit s0 + 2 bit s1.
// ---------------------------------------------------------------------------
template <typename T, int NV>
__device__ __forceinline__ void select_block(T (&out)[NV], CT<T> *base, int nb, bool c0, bool c1) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
        const T x0 = base[i], x1 = base[NV + i];
        T lo = c0 ? x1 : x0;
        if (nb == 4) {
            const T x2 = base[2 * NV + i], x3 = base[3 * NV + i];
            const T hi = c0 ? x3 : x2;
            lo = c1 ? hi : lo;
        }
        out[i] = lo;
    }
}

// ---------------------------------------------------------------------------
// Out-of-place operations on one state: dst = U src.
// ---------------------------------------------------------------------------
template <typename T, int A, int Q>
__device__ __forceinline__ void u1_diag(Cx<T> (&d)[A], const Cx<T> (&s)[A], Cx<T> u0, Cx<T> u1) {
#pragma unroll
    for (int k = 0; k < A; k++) {
        d[k] = cmul((k >> Q) & 1 ? u1 : u0, s[k]);
    }
}

template <typename T, int A, int Q>
__device__ __forceinline__ void u1_real(Cx<T> (&d)[A], const Cx<T> (&s)[A], T m00, T m01, T m10,
                                        T m11) {
#pragma unroll
    for (int k = 0; k < A; k++) {
        if (k & (1 << Q)) {
            continue;
        }
        const Cx<T> a = s[k], b = s[k | (1 << Q)];
        d[k] = {fma(m00, a.r, m01 * b.r), fma(m00, a.i, m01 * b.i)};
        d[k | (1 << Q)] = {fma(m10, a.r, m11 * b.r), fma(m10, a.i, m11 * b.i)};
    }
}

template <typename T, int A, int Q>
__device__ __forceinline__ void u1_dense(Cx<T> (&d)[A], const Cx<T> (&s)[A], Cx<T> u00, Cx<T> u01,
                                         Cx<T> u10, Cx<T> u11) {
#pragma unroll
    for (int k = 0; k < A; k++) {
        if (k & (1 << Q)) {
            continue;
        }
        const Cx<T> a = s[k], b = s[k | (1 << Q)];
        d[k] = cmad(u01, b, cmul(u00, a));
        d[k | (1 << Q)] = cmad(u11, b, cmul(u10, a));
    }
}

// The per-lane condition selects between values, never between array
// elements (a select of two element addresses would put the whole register
// tile in scratch memory).
template <typename T, int A, int Q>
__device__ __forceinline__ void u1_swap(Cx<T> (&d)[A], const Cx<T> (&s)[A], bool on) {
#pragma unroll
    for (int k = 0; k < A; k++) {
        const Cx<T> x = s[k ^ (1 << Q)], y = s[k];
        d[k] = {on ? x.r : y.r, on ? x.i : y.i};
    }
}

template <int QA, int QB> __device__ __forceinline__ constexpr int lidx(int k) {
    return (((k >> QA) & 1) << 1) | ((k >> QB) & 1);
}

template <typename T, int A, int QA, int QB>
__device__ __forceinline__ void u2_dense(Cx<T> (&d)[A], const Cx<T> (&s)[A], CT<T> *U) {
    constexpr int MQ = (1 << QA) | (1 << QB);
#pragma unroll
    for (int k = 0; k < A; k++) {
        if (k & MQ) {
            continue;
        }
        const int idx[4] = {k, k | (1 << QB), k | (1 << QA), k | MQ};
#pragma unroll
        for (int r = 0; r < 4; r++) {
            Cx<T> t = cmul(Cx<T>{U[8 * r], U[8 * r + 1]}, s[idx[0]]);
#pragma unroll
            for (int c = 1; c < 4; c++) {
                t = cmad(Cx<T>{U[8 * r + 2 * c], U[8 * r + 2 * c + 1]}, s[idx[c]], t);
            }
            d[idx[r]] = t;
        }
    }
}

template <typename T, int A, int QA, int QB>
__device__ __forceinline__ void u2_real(Cx<T> (&d)[A], const Cx<T> (&s)[A], CT<T> *U) {
    constexpr int MQ = (1 << QA) | (1 << QB);
#pragma unroll
    for (int k = 0; k < A; k++) {
        if (k & MQ) {
            continue;
        }
        const int idx[4] = {k, k | (1 << QB), k | (1 << QA), k | MQ};
#pragma unroll
        for (int r = 0; r < 4; r++) {
            T sr = U[4 * r] * s[idx[0]].r, si = U[4 * r] * s[idx[0]].i;
#pragma unroll
            for (int c = 1; c < 4; c++) {
                sr = fma(U[4 * r + c], s[idx[c]].r, sr);
                si = fma(U[4 * r + c], s[idx[c]].i, si);
            }
            d[idx[r]] = {sr, si};
        }
    }
}

template <typename T, int A, int QA, int QB>
__device__ __forceinline__ void u2_diag(Cx<T> (&d)[A], const Cx<T> (&s)[A], CT<T> *U) {
#pragma unroll
    for (int k = 0; k < A; k++) {
        const int r = lidx<QA, QB>(k);
        d[k] = cmul(Cx<T>{U[2 * r], U[2 * r + 1]}, s[k]);
    }
}

// Local indices LA and LB exchanged, every other amplitude copied.
template <typename T, int A, int QA, int QB, int LA, int LB>
__device__ __forceinline__ void u2_swap(Cx<T> (&d)[A], const Cx<T> (&s)[A]) {
    constexpr int flip = ((((LA ^ LB) >> 1) & 1) << QA) | (((LA ^ LB) & 1) << QB);
#pragma unroll
    for (int k = 0; k < A; k++) {
        const int r = lidx<QA, QB>(k);
        d[k] = (r == LA || r == LB) ? s[k ^ flip] : s[k];
    }
}

// gf = G f for a one-register-bit generator.
template <typename T, int A, int Q>
__device__ __forceinline__ void gen1(Cx<T> (&gf)[A], const Cx<T> (&f)[A], int gkind, CT<T> *G) {
    if (gkind == G_Z) {
#pragma unroll
        for (int k = 0; k < A; k++) {
            gf[k] = ((k >> Q) & 1) ? Cx<T>{-f[k].r, -f[k].i} : f[k];
        }
    } else if (gkind == G_X) {
#pragma unroll
        for (int k = 0; k < A; k++) {
            gf[k] = f[k ^ (1 << Q)];
        }
    } else if (gkind == G_Y) {
        // Y = [[0, -i], [i, 0]]: (Y f)_0 = -i f_1, (Y f)_1 = i f_0.
#pragma unroll
        for (int k = 0; k < A; k++) {
            const Cx<T> p = f[k ^ (1 << Q)];
            gf[k] = ((k >> Q) & 1) ? Cx<T>{-p.i, p.r} : Cx<T>{p.i, -p.r};
        }
    } else if (gkind == G_DIAG) {
        const T g0 = G[0], g1 = G[1];
#pragma unroll
        for (int k = 0; k < A; k++) {
            const T g = ((k >> Q) & 1) ? g1 : g0;
            gf[k] = {g * f[k].r, g * f[k].i};
        }
    } else {
        u1_dense<T, A, Q>(gf, f, {G[0], G[1]}, {G[2], G[3]}, {G[4], G[5]}, {G[6], G[7]});
    }
}

template <typename T, int A, int QA, int QB>
__device__ __forceinline__ void gen2(Cx<T> (&gf)[A], const Cx<T> (&f)[A], int gkind, CT<T> *G) {
    if (gkind == G_Z) {
#pragma unroll
        for (int k = 0; k < A; k++) {
            gf[k] = (__builtin_popcount(lidx<QA, QB>(k)) & 1) ? Cx<T>{-f[k].r, -f[k].i} : f[k];
        }
    } else if (gkind == G_DIAG) {
#pragma unroll
        for (int k = 0; k < A; k++) {
            const T g = G[lidx<QA, QB>(k)];
            gf[k] = {g * f[k].r, g * f[k].i};
        }
    } else {
        u2_dense<T, A, QA, QB>(gf, f, G);
    }
}

// ---------------------------------------------------------------------------
// Inner products of a trainable operation: d_j = scale sum_k Im(conj(h_j[k])
// gf[k]) in FP64, summed over the quad by DPP and added into LDS by the
// quad's first lane (address (slot, j, quad)).
// ---------------------------------------------------------------------------
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b & 0xffffffffu), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(static_cast<long long>(
        (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo)));
}

template <typename T, int NS, int A, int ROT>
__device__ __forceinline__ void accumulate(const Cx<T> (&B)[NS + 1][A], const Cx<T> (&gf)[A],
                                           double scale, double *lacc, int slot, int lane) {
    using RT = Rot<NS, ROT>;
    constexpr int NJ = NS - 1;
    double d[NJ > 0 ? NJ : 1];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < A; k++) {
            s += im_conj_mul(B[RT::in(j + 1)][k], gf[k]);
        }
        s *= scale;
        s += dpp_f64<0xB1>(s); // quad_perm [1,0,3,2]
        s += dpp_f64<0x4E>(s); // quad_perm [2,3,0,1]
        d[j] = s;
    }
    // Lane r of a quad adds the totals of observables r and r + 4 (every lane
    // holds every total); lanes without one add to a dummy slot. No branch on
    // the lane, so the instruction switch around this stays free of divergent
    // control flow.
    const int r = lane & 3;
#pragma unroll
    for (int j0 = 0; j0 < NJ; j0 += 4) {
        double v = d[j0];
#pragma unroll
        for (int j = j0 + 1; j < j0 + 4 && j < NJ; j++) {
            v = (r == j - j0) ? d[j] : v;
        }
        const int j = j0 + r;
        const int idx = j < NJ ? (slot * NJ + j) * 16 + (lane >> 2) : RT_PMAX * NJ * 16 + lane;
        atomicAdd(&lacc[idx], v);
    }
}

// ---------------------------------------------------------------------------
// One instruction at rotation ROT.
// ---------------------------------------------------------------------------
template <int R, typename F> __device__ __forceinline__ void dispatch_q(int q, F &&f) {
    switch (q) {
    case 0: f(IC<0>{}); break;
    case 1: f(IC<1>{}); break;
    case 2: f(IC<2>{}); break;
    default:
        if constexpr (R > 3) {
            f(IC<3>{});
        }
        break;
    }
}

template <int R, typename F> __device__ __forceinline__ void dispatch_q2(int a, int b, F &&f) {
    switch (a * 4 + b) {
    case 4: f(IC<1>{}, IC<0>{}); break;
    case 8: f(IC<2>{}, IC<0>{}); break;
    case 9: f(IC<2>{}, IC<1>{}); break;
    default:
        if constexpr (R > 3) {
            switch (a * 4 + b) {
            case 12: f(IC<3>{}, IC<0>{}); break;
            case 13: f(IC<3>{}, IC<1>{}); break;
            default: f(IC<3>{}, IC<2>{}); break;
            }
        }
        break;
    }
}

template <typename T, int NS, int A, int R, int ROT>
__device__ __forceinline__ void remap_all(Cx<T> (&B)[NS + 1][A], int q, int p) {
    using RT = Rot<NS, ROT>;
    dispatch_q<R>(q, [&](auto Qc) {
        constexpr int Q = decltype(Qc)::value;
        if (p == 5) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                remap1<T, A, Q, 5>(B[RT::in(s)]);
            }
        } else {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                remap1<T, A, Q, 4>(B[RT::in(s)]);
            }
        }
    });
}

template <typename T, int NS, int A, int R, int ROT>
__device__ __forceinline__ void step(Cx<T> (&B)[NS + 1][A], CT<unsigned char> *prog, int pc,
                                     double *lacc, int lane, uint64_t tbase) {
    using RT = Rot<NS, ROT>;
    const auto *words = reinterpret_cast<CT<int32_t> *>(prog) + 8 * pc;
    int32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        w[k] = words[k];
    }
    const int code = static_cast<int8_t>(w[0]);
    const int a = static_cast<int8_t>(w[0] >> 8);
    const int b = static_cast<int8_t>(w[0] >> 16);
    const int s0 = static_cast<int8_t>(w[0] >> 24);
    const int s1 = static_cast<int8_t>(w[1]);
    const int ukind = static_cast<int8_t>(w[1] >> 8);
    const int gkind = static_cast<int8_t>(w[1] >> 16);
    const int swp = static_cast<int8_t>(w[1] >> 24);
    const int slot = w[2];
    CT<T> *U = reinterpret_cast<CT<T> *>(prog + w[3]);
    CT<T> *G = reinterpret_cast<CT<T> *>(prog + w[4]);
    const int nremap = static_cast<int8_t>(w[5]);
    if (nremap > 0) {
        remap_all<T, NS, A, R, ROT>(B, static_cast<int8_t>(w[5] >> 8), static_cast<int8_t>(w[5] >> 16));
        if (nremap > 1) {
            remap_all<T, NS, A, R, ROT>(B, static_cast<int8_t>(w[5] >> 24), static_cast<int8_t>(w[6]));
        }
    }
    auto &gf = B[RT::out(0)]; // free until state 0 is written
    const auto &phi = B[RT::in(0)];
    if (code == RT_REG1) {
        dispatch_q<R>(a, [&](auto Qc) {
            constexpr int Q = decltype(Qc)::value;
            if (slot >= 0) {
                gen1<T, A, Q>(gf, phi, gkind, G);
                accumulate<T, NS, A, ROT>(B, gf, 1.0, lacc, slot, lane);
            }
            if (ukind == U_DIAG) {
                const Cx<T> u0{U[0], U[1]}, u1{U[2], U[3]};
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u1_diag<T, A, Q>(B[RT::out(s)], B[RT::in(s)], u0, u1);
                }
            } else if (ukind == U_REAL) {
                const T m00 = U[0], m01 = U[1], m10 = U[2], m11 = U[3];
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u1_real<T, A, Q>(B[RT::out(s)], B[RT::in(s)], m00, m01, m10, m11);
                }
            } else if (ukind == U_SWAP) {
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u1_swap<T, A, Q>(B[RT::out(s)], B[RT::in(s)], true);
                }
            } else {
                const Cx<T> u00{U[0], U[1]}, u01{U[2], U[3]}, u10{U[4], U[5]}, u11{U[6], U[7]};
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u1_dense<T, A, Q>(B[RT::out(s)], B[RT::in(s)], u00, u01, u10, u11);
                }
            }
        });
    } else if (code == RT_REG2) {
        dispatch_q2<R>(a, b, [&](auto QAc, auto QBc) {
            constexpr int QA = decltype(QAc)::value, QB = decltype(QBc)::value;
            if (slot >= 0) {
                gen2<T, A, QA, QB>(gf, phi, gkind, G);
                accumulate<T, NS, A, ROT>(B, gf, 1.0, lacc, slot, lane);
            }
            if (ukind == U_DIAG) {
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u2_diag<T, A, QA, QB>(B[RT::out(s)], B[RT::in(s)], U);
                }
            } else if (ukind == U_SWAP) {
                auto sw = [&](auto LA, auto LB) {
#pragma unroll
                    for (int s = 0; s < NS; s++) {
                        u2_swap<T, A, QA, QB, decltype(LA)::value, decltype(LB)::value>(B[RT::out(s)],
                                                                                        B[RT::in(s)]);
                    }
                };
                switch (swp) {
                case 1: sw(IC<0>{}, IC<1>{}); break;
                case 2: sw(IC<0>{}, IC<2>{}); break;
                case 3: sw(IC<0>{}, IC<3>{}); break;
                case 6: sw(IC<1>{}, IC<2>{}); break;
                case 7: sw(IC<1>{}, IC<3>{}); break;
                default: sw(IC<2>{}, IC<3>{}); break;
                }
            } else if (ukind == U_REAL) {
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u2_real<T, A, QA, QB>(B[RT::out(s)], B[RT::in(s)], U);
                }
            } else {
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u2_dense<T, A, QA, QB>(B[RT::out(s)], B[RT::in(s)], U);
                }
            }
        });
    } else if (code == RT_SEL) {
        const int nb = s1 >= 0 ? 4 : 2;
        const bool c0 = selbit(lane, tbase, s0);
        const bool c1 = s1 >= 0 ? selbit(lane, tbase, s1) : false;
        dispatch_q<R>(a, [&](auto Qc) {
            constexpr int Q = decltype(Qc)::value;
            if (slot >= 0) {
                if (gkind == G_DIAG) {
                    T g[2];
                    select_block<T, 2>(g, G, nb, c0, c1);
#pragma unroll
                    for (int k = 0; k < A; k++) {
                        const T x = ((k >> Q) & 1) ? g[1] : g[0];
                        gf[k] = {x * phi[k].r, x * phi[k].i};
                    }
                } else {
                    T g[8];
                    select_block<T, 8>(g, G, nb, c0, c1);
                    u1_dense<T, A, Q>(gf, phi, {g[0], g[1]}, {g[2], g[3]}, {g[4], g[5]}, {g[6], g[7]});
                }
                accumulate<T, NS, A, ROT>(B, gf, 1.0, lacc, slot, lane);
            }
            if (ukind == U_DIAG) {
                T u[4];
                select_block<T, 4>(u, U, nb, c0, c1);
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u1_diag<T, A, Q>(B[RT::out(s)], B[RT::in(s)], {u[0], u[1]}, {u[2], u[3]});
                }
            } else if (ukind == U_SWAP) {
                T fl[1];
                select_block<T, 1>(fl, U, nb, c0, c1);
                const bool on = fl[0] != T(0);
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u1_swap<T, A, Q>(B[RT::out(s)], B[RT::in(s)], on);
                }
            } else if (ukind == U_REAL) {
                T u[4];
                select_block<T, 4>(u, U, nb, c0, c1);
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u1_real<T, A, Q>(B[RT::out(s)], B[RT::in(s)], u[0], u[1], u[2], u[3]);
                }
            } else {
                T u[8];
                select_block<T, 8>(u, U, nb, c0, c1);
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    u1_dense<T, A, Q>(B[RT::out(s)], B[RT::in(s)], {u[0], u[1]}, {u[2], u[3]},
                                      {u[4], u[5]}, {u[6], u[7]});
                }
            }
        });
    } else {
        // RT_LANE: G phi = g phi on this lane, so the products scale by g.
        const int nb = s1 >= 0 ? 4 : 2;
        const bool c0 = selbit(lane, tbase, s0);
        const bool c1 = s1 >= 0 ? selbit(lane, tbase, s1) : false;
        if (slot >= 0) {
            T g[1];
            select_block<T, 1>(g, G, nb, c0, c1);
            accumulate<T, NS, A, ROT>(B, phi, static_cast<double>(g[0]), lacc, slot, lane);
        }
        T u[2];
        select_block<T, 2>(u, U, nb, c0, c1);
        const Cx<T> f{u[0], u[1]};
#pragma unroll
        for (int s = 0; s < NS; s++) {
#pragma unroll
            for (int k = 0; k < A; k++) {
                B[RT::out(s)][k] = cmul(f, B[RT::in(s)][k]);
            }
        }
    }
}

// Runs the program with the loop unrolled over the NS + 1 rotations; returns
// the rotation after the last instruction.
template <typename T, int NS, int A, int R, int... ROTS>
__device__ __forceinline__ int run_program(Cx<T> (&B)[NS + 1][A], CT<unsigned char> *prog, int ninstr,
                                           uint64_t tbase,
                                           double *lacc, int lane,
                                           std::integer_sequence<int, ROTS...>) {
    int pc = 0;
    for (;;) {
        int end = 0;
        const bool more = ((pc < ninstr ? (step<T, NS, A, R, ROTS>(B, prog, pc, lacc, lane, tbase), ++pc, true)
                                        : (end = ROTS, false)) &&
                           ...);
        if (!more) {
            return end;
        }
    }
}

template <typename F, int... RS>
__device__ __forceinline__ void for_rot(int rot, F &&f, std::integer_sequence<int, RS...>) {
    ((rot == RS ? (f(IC<RS>{}), 0) : 0), ...);
}

// Observable-state pointers of one launch: nc chunks of NJ states.
template <typename T> struct RtPtrs {
    Cx<T> *p[RT_OBS_MAX];
};

template <typename T, int NJ, typename Off>
__global__ void __launch_bounds__(64)
    rtile_kernel(Cx<T> *phi, Cx<T> *phi_out, RtPtrs<T> hp, const unsigned char *__restrict__ prog_g,
                 RtLaunch L, double *__restrict__ partials) {
    constexpr int R = rt_r<T>();
    constexpr int A = 1 << R;
    constexpr int NS = NJ + 1;
    constexpr int NA = NJ > 0 ? NJ : 1;
    // Block -> (chunk c, block b of that chunk). Ids are taken in groups of
    // xgroup (consecutive ids land on different XCDs), so the nc blocks that
    // stream the same tiles for different chunks share an XCD and its L2.
    const int nc = L.nc;
    const unsigned xg = static_cast<unsigned>(L.xgroup);
    const unsigned kq = blockIdx.x / xg;
    const int c = static_cast<int>(kq % static_cast<unsigned>(nc));
    const unsigned b = (kq / static_cast<unsigned>(nc)) * xg + blockIdx.x % xg;
    const unsigned bpc = gridDim.x / static_cast<unsigned>(nc);
    const bool writes_phi = L.write_phi && c == nc - 1;
    // This chunk's states and its first state in the launch: nbig chunks of
    // NJ, then chunks of NJ - 1 (their last register state stays unused).
    const int njc = c < L.nbig ? NJ : NJ - 1;
    const int jc0 = c < L.nbig ? c * NJ : L.nbig * NJ + (c - L.nbig) * (NJ - 1);
    const int width = L.nbig * NJ + (nc - L.nbig) * (NJ - 1);
    __shared__ double lacc[NJ > 0 ? RT_PMAX * NA * 16 + 64 : 1]; // + one dummy slot per lane
    const int lane = static_cast<int>(__lane_id());
    if constexpr (NJ > 0) {
        const int nacc = L.nslots * NA * 16;
        for (int i = lane; i < nacc; i += 64) {
            lacc[i] = 0.0;
        }
    }
    CT<unsigned char> *prog = (CT<unsigned char> *)(prog_g);
    Cx<T> *sp[NS];
    sp[0] = phi;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        sp[j + 1] = j < njc ? hp.p[jc0 + j] : phi;
    }
    // Per-lane offsets of the load and store layouts (amplitudes; the tile
    // base is added per tile); register slot k sits at the sum over its set
    // bits q of 2^reg[q] (uniform).
    uint64_t lin = 0, lout = 0;
#pragma unroll
    for (int p = 0; p < RT_LANE_BITS; p++) {
        lin |= static_cast<uint64_t>(lbit(lane, p)) << L.lane_in[p];
        lout |= static_cast<uint64_t>(lbit(lane, p)) << L.lane_out[p];
    }
    const Off vin = static_cast<Off>(lin * sizeof(Cx<T>));
    const Off vout = static_cast<Off>(lout * sizeof(Cx<T>));
    auto roff = [&](const int *reg, int k) {
        uint64_t o = 0;
#pragma unroll
        for (int q = 0; q < R; q++) {
            if ((k >> q) & 1) {
                o += uint64_t{1} << reg[q];
            }
        }
        return o;
    };
    const uint64_t ntiles = uint64_t{1} << (L.n - L.T);
    for (uint64_t g = b; g < ntiles; g += bpc) {
        uint64_t base = g;
        for (int q = 0; q < L.T; q++) {
            base = insert_zero(base, L.tb[q]);
        }
        Cx<T> B[NS + 1][A];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            if (s > njc) {
#pragma unroll
                for (int k = 0; k < A; k++) {
                    B[s][k] = Cx<T>{T(0), T(0)};
                }
                continue;
            }
#pragma unroll
            for (int k = 0; k < A; k++) {
                B[s][k] = ldo(sp[s], static_cast<Off>((base + roff(L.reg_in, k)) * sizeof(Cx<T>)) + vin);
            }
        }
        const int rot = run_program<T, NS, A, R>(B, prog, L.ninstr, base, lacc, lane,
                                                 std::make_integer_sequence<int, NS + 1>{});
        auto store = [&](auto RC) {
            using RT = Rot<NS, decltype(RC)::value>;
            // The store layout's remaps (low index bits back onto lanes).
            for (int f = 0; f < L.nfin; f++) {
                remap_all<T, NS, A, R, decltype(RC)::value>(B, L.fin_rq[f], L.fin_rp[f]);
            }
#pragma unroll
            for (int s = 0; s < NS; s++) {
                if ((s == 0 && !writes_phi) || s > njc) {
                    continue;
                }
                Cx<T> *dst = s == 0 ? phi_out : sp[s];
#pragma unroll
                for (int k = 0; k < A; k++) {
                    sto(dst, static_cast<Off>((base + roff(L.reg_out, k)) * sizeof(Cx<T>)) + vout,
                        B[RT::in(s)][k]);
                }
            }
        };
        for_rot(rot, store, std::make_integer_sequence<int, NS + 1>{});
    }
    if constexpr (NJ > 0) {
        __syncthreads();
        const int per = L.nslots * NJ;
        for (int i = lane; i < per; i += 64) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 16; q++) {
                s += lacc[i * 16 + q];
            }
            // Local (slot, j) -> launch-wide (slot, jc0 + j).
            const int slot = i / NJ, j = i % NJ;
            if (j < njc) {
                const uint64_t row = static_cast<uint64_t>(slot) * width + jc0 + j;
                partials[row * bpc + b] = s;
            }
        }
    }
}

} // namespace

// Launcher of one (precision, observable chunk) variant. The build compiles
// this file once per variant with PLK_RT_T and PLK_RT_NJ defined, so the
// variants compile in parallel.
template <typename T, int NJ>
int rtile_launch_t(Device &dev, const State &phi, void *phi_out, void *const *h_ptrs,
                   const void *d_prog, const RtLaunch &Lin) {
    RtLaunch L = Lin;
    L.nc = std::max(1, L.nc);
    L.nbig = std::min(std::max(L.nbig, NJ > 0 ? 1 : 0), L.nc);
    if (NJ == 0) {
        L.nbig = L.nc;
    }
    const int width = L.nbig * NJ + (L.nc - L.nbig) * std::max(0, NJ - 1);
    RtPtrs<T> hp{};
    for (int j = 0; j < width; j++) {
        hp.p[j] = static_cast<Cx<T> *>(h_ptrs[j]);
    }
    const void *kfn = reinterpret_cast<const void *>(&rtile_kernel<T, NJ, uint64_t>);
    static int blocks_per_cu = 0;
    if (blocks_per_cu == 0) {
        int nb = 0;
        PLK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kfn, 64, 0));
        blocks_per_cu = std::max(1, std::min(nb, 32));
    }
    const uint64_t ntiles = uint64_t{1} << (L.n - L.T);
    const uint64_t cap = static_cast<uint64_t>(dev.num_cus) * blocks_per_cu;
    // Blocks per chunk: the chunks share the resident blocks; in groups of 8
    // (one per XCD) when there are enough.
    uint64_t bpc = std::max<uint64_t>(1, std::min(ntiles, cap / static_cast<uint64_t>(L.nc)));
    L.xgroup = 1;
    if (bpc >= 8) {
        bpc = bpc / 8 * 8;
        L.xgroup = 8;
    }
    const unsigned grid = static_cast<unsigned>(bpc * static_cast<uint64_t>(L.nc));
    if (NJ > 0 && L.nslots > 0) {
        ensure_red(dev, static_cast<size_t>(bpc) * L.nslots * width, 0);
    }
    // Algorithmic bytes: phi read once and written once, every observable
    // state read and written once.
    const double bytes = (1.0 + L.write_phi + 2.0 * width) * static_cast<double>(phi.bytes());
    ProfScope prof(NJ > 0 ? "adjoint_rtile" : "apply_rtile", bytes, dev.stream);
    hipLaunchKernelGGL((rtile_kernel<T, NJ, uint64_t>), dim3(grid), dim3(64), 0, dev.stream,
                       static_cast<Cx<T> *>(phi.d), static_cast<Cx<T> *>(phi_out), hp,
                       static_cast<const unsigned char *>(d_prog), L, dev.red);
    PLK_HIP(hipGetLastError());
    return static_cast<int>(bpc);
}

#if defined(PLK_RT_T) && defined(PLK_RT_NJ)
template int rtile_launch_t<PLK_RT_T, PLK_RT_NJ>(Device &, const State &, void *, void *const *,
                                                 const void *, const RtLaunch &);
#endif

} // namespace plk
