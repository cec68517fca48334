// Single-operation adjoint steps for the operations the fused passes do not
// hold: dense matrices on 3 or 4 wires (Toffoli, CSWAP, DoubleExcitation±,
// 3-4-wire matrices, MultiRZ on 3-4 wires) and parity phases on any number of
// wires (MultiRZ).
//
// The reference takes one such step as a copy mu = lambda, U^dagger on
// lambda, G on mu, one blocking inner product per observable and U^dagger on
// every H_j lambda (AdjointDiffKokkos.hpp:454-476): 3 + 2 n_obs launches.
// Here one launch per chunk of observable states does the whole step. An op
// on m wires acts on groups of 2^m amplitudes (the op's wires, all other
// bits fixed), but most such ops move only a few of a group's states
// (DoubleExcitation: 2 of 16, Toffoli and CSWAP: 2 of 8). The host passes
// the D-state subspace outside which U is the identity and G is zero
// (D = 2, 4, 8 or 16) with U and G restricted to it. Each thread owns one
// group: it loads phi's D amplitudes, forms G phi in registers (phi before
// U, as the reference's mu), stores U phi (with the last chunk), and for each
// observable state h_j loads the group's D amplitudes, adds Im<h_j|G phi> in
// FP64 and writes U h_j back: (1 + w + 2 nj) (D / 2^m) S per launch, in
// whole cache lines when the op's wires include the lowest index bits. U and
// G are FP64 matrices read through the scalar cache; amplitudes of complex64
// states are widened to FP64 for the products and rounded once on the
// store.
//
// Jacobian sums: per thread in LDS (one slot per observable state), then a
// fixed-order block sum into dev.red as [j][block]; tile_reduce_kernel adds
// the blocks in order.
#include "device_common.hpp"

namespace plk {

namespace {

constexpr int kStepBlock = 256;

__device__ __forceinline__ double block_total(double v, double *red) {
    v = wave_sum(v);
    __syncthreads(); // red is reused across calls
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// Writes the block's FP64 totals of the nj per-thread slots.
__device__ __forceinline__ void step_partials(const double *sacc, int nj, double *partials) {
    __shared__ double red[kStepBlock / 64];
    for (int j = 0; j < nj; j++) {
        const double t = block_total(sacc[j * kStepBlock + threadIdx.x], red);
        if (threadIdx.x == 0) {
            partials[static_cast<uint64_t>(j) * gridDim.x + blockIdx.x] = t;
        }
    }
}

// w = M x for a D x D FP64 matrix (row-major, double2 entries); stored
// through the group's offsets.
template <typename T, int D>
__device__ __forceinline__ void mat_store(Cx<T> *p, uint64_t base, const uint64_t *off,
                                          const double2 *__restrict__ M, const double *xr,
                                          const double *xi) {
#pragma unroll
    for (int r = 0; r < D; r++) {
        // One row of matrix loads at a time (scalar registers hold a row).
        __asm__ volatile("" ::: "memory");
        double wr = 0.0, wi = 0.0;
#pragma unroll
        for (int c = 0; c < D; c++) {
            const double2 m = M[r * D + c];
            wr = fma(m.x, xr[c], fma(-m.y, xi[c], wr));
            wi = fma(m.x, xi[c], fma(m.y, xr[c], wi));
        }
        st(p, base + off[r], Cx<T>{static_cast<T>(wr), static_cast<T>(wi)});
    }
}

template <typename T, int D>
__global__ void __launch_bounds__(kStepBlock)
    step_dense_kernel(Cx<T> *__restrict__ phi, SegPtrs<T> hp, const double2 *__restrict__ mats,
                      StepLaunch a, double *__restrict__ partials) {
    const double2 *U = mats;
    const double2 *G = mats + D * D;
    __shared__ double sacc[SEG_NJMAX * kStepBlock];
    for (int j = 0; j < a.nj; j++) {
        sacc[j * kStepBlock + threadIdx.x] = 0.0;
    }
    uint64_t off[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        off[k] = a.off[k];
    }
    for (uint64_t g = uint64_t{blockIdx.x} * kStepBlock + threadIdx.x; g < a.groups;
         g += uint64_t{gridDim.x} * kStepBlock) {
        uint64_t base = g;
#pragma unroll
        for (int q = 0; q < STEP_MMAX; q++) {
            if (q < a.m) {
                base = insert_zero(base, a.bits[q]);
            }
        }
        double vr[D], vi[D];
#pragma unroll
        for (int k = 0; k < D; k++) {
            const Cx<T> c = ld(phi, base + off[k]);
            vr[k] = c.r;
            vi[k] = c.i;
        }
        double gr[D], gi[D];
        if (a.trainable) {
#pragma unroll
            for (int r = 0; r < D; r++) {
                __asm__ volatile("" ::: "memory");
                double wr = 0.0, wi = 0.0;
#pragma unroll
                for (int c = 0; c < D; c++) {
                    const double2 m = G[r * D + c];
                    wr = fma(m.x, vr[c], fma(-m.y, vi[c], wr));
                    wi = fma(m.x, vi[c], fma(m.y, vr[c], wi));
                }
                gr[r] = wr;
                gi[r] = wi;
            }
        }
        // phi is read only here, so its update can be stored first and its
        // registers reused for the observable states.
        if (a.write_phi) {
            mat_store<T, D>(phi, base, off, U, vr, vi);
        }
        for (int j = 0; j < a.nj; j++) {
            Cx<T> *h = hp.p[j];
            double xr[D], xi[D];
#pragma unroll
            for (int k = 0; k < D; k++) {
                const Cx<T> c = ld(h, base + off[k]);
                xr[k] = c.r;
                xi[k] = c.i;
            }
            if (a.trainable) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < D; k++) {
                    s = fma(xr[k], gi[k], fma(-xi[k], gr[k], s));
                }
                sacc[j * kStepBlock + threadIdx.x] += s;
            }
            mat_store<T, D>(h, base, off, U, xr, xi);
        }
    }
    if (a.trainable) {
        step_partials(sacc, a.nj, partials);
    }
}

// Parity phases: U multiplies amplitude i by ue or uo by the parity of
// i & umask; G multiplies it by +1 / -1 by the parity of i & gmask.
template <typename T>
__global__ void __launch_bounds__(kStepBlock)
    step_parity_kernel(Cx<T> *__restrict__ phi, SegPtrs<T> hp, StepLaunch a,
                       double *__restrict__ partials) {
    __shared__ double sacc[SEG_NJMAX * kStepBlock];
    for (int j = 0; j < a.nj; j++) {
        sacc[j * kStepBlock + threadIdx.x] = 0.0;
    }
    for (uint64_t i = uint64_t{blockIdx.x} * kStepBlock + threadIdx.x; i < a.groups;
         i += uint64_t{gridDim.x} * kStepBlock) {
        const bool uodd = __popcll(i & a.umask) & 1;
        const double pr = uodd ? a.uo[0] : a.ue[0];
        const double pi = uodd ? a.uo[1] : a.ue[1];
        const double sg = (__popcll(i & a.gmask) & 1) ? -1.0 : 1.0;
        const Cx<T> v = ld(phi, i);
        const double gr = sg * static_cast<double>(v.r), gi = sg * static_cast<double>(v.i);
        for (int j = 0; j < a.nj; j++) {
            Cx<T> *h = hp.p[j];
            const Cx<T> x = ld(h, i);
            const double xr = x.r, xi = x.i;
            if (a.trainable) {
                sacc[j * kStepBlock + threadIdx.x] += fma(xr, gi, -xi * gr);
            }
            st(h, i, Cx<T>{static_cast<T>(pr * xr - pi * xi), static_cast<T>(pr * xi + pi * xr)});
        }
        if (a.write_phi) {
            const double xr = v.r, xi = v.i;
            st(phi, i, Cx<T>{static_cast<T>(pr * xr - pi * xi), static_cast<T>(pr * xi + pi * xr)});
        }
    }
    if (a.trainable) {
        step_partials(sacc, a.nj, partials);
    }
}

template <typename T>
int launch_step_t(Device &dev, const State &phi, void *const *h_ptrs, const void *d_mats,
                  const StepLaunch &L) {
    SegPtrs<T> hp{};
    for (int j = 0; j < L.nj; j++) {
        hp.p[j] = static_cast<Cx<T> *>(h_ptrs[j]);
    }
    const uint64_t blocks = (L.groups + kStepBlock - 1) / kStepBlock;
    const unsigned grid = static_cast<unsigned>(
        std::max<uint64_t>(1, std::min<uint64_t>(blocks, static_cast<uint64_t>(dev.num_cus) * 8)));
    if (L.trainable && L.nj > 0) {
        ensure_red(dev, static_cast<size_t>(grid) * L.nj, 0);
    }
    const double bytes = (1.0 + L.write_phi + 2.0 * L.nj) * static_cast<double>(phi.bytes());
    ProfScope prof(L.nj > 0 ? "adjoint_step" : "apply_step", bytes, dev.stream);
    auto *p = static_cast<Cx<T> *>(phi.d);
    const auto *m = static_cast<const double2 *>(d_mats);
    switch (L.m == 0 ? 0 : L.dim) {
    case 0:
        hipLaunchKernelGGL((step_parity_kernel<T>), dim3(grid), dim3(kStepBlock), 0, dev.stream, p,
                           hp, L, dev.red);
        break;
    case 2:
        hipLaunchKernelGGL((step_dense_kernel<T, 2>), dim3(grid), dim3(kStepBlock), 0, dev.stream,
                           p, hp, m, L, dev.red);
        break;
    case 4:
        hipLaunchKernelGGL((step_dense_kernel<T, 4>), dim3(grid), dim3(kStepBlock), 0, dev.stream,
                           p, hp, m, L, dev.red);
        break;
    case 8:
        hipLaunchKernelGGL((step_dense_kernel<T, 8>), dim3(grid), dim3(kStepBlock), 0, dev.stream,
                           p, hp, m, L, dev.red);
        break;
    case 16:
        hipLaunchKernelGGL((step_dense_kernel<T, 16>), dim3(grid), dim3(kStepBlock), 0, dev.stream,
                           p, hp, m, L, dev.red);
        break;
    default:
        raise(PLK_ERR_INVALID, "single-op step kernels take subspaces of 2, 4, 8 or 16 states");
    }
    PLK_HIP(hipGetLastError());
    return static_cast<int>(grid);
}

} // namespace

int launch_step(Device &dev, const State &phi, void *const *h_ptrs, const void *d_mats,
                const StepLaunch &L) {
    if (L.nj < 0 || L.nj > SEG_NJMAX) {
        raise(PLK_ERR_INVALID, "single-op step: observable chunk out of range");
    }
    if (L.m != 0 && (L.m < 1 || L.m > STEP_MMAX || d_mats == nullptr || L.dim < 2 ||
                     L.dim > (1 << L.m))) {
        raise(PLK_ERR_INVALID, "single-op step: bad matrix width");
    }
    if (L.m > 0 && (uint64_t{1} << L.m) * L.groups != phi.len) {
        raise(PLK_ERR_INVALID, "single-op step: groups do not cover the state");
    }
    if (L.m == 0 && L.groups != phi.len) {
        raise(PLK_ERR_INVALID, "single-op step: parity step must cover the state");
    }
    return phi.dtype == PLK_C64 ? launch_step_t<float>(dev, phi, h_ptrs, d_mats, L)
                                : launch_step_t<double>(dev, phi, h_ptrs, d_mats, L);
}

} // namespace plk
