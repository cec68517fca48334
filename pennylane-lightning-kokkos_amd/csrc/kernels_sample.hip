// Probabilities and sampling (MeasuresKokkos.hpp:389-569,
// MeasuresFunctors.hpp:15-210), FP64 and deterministic.
//
// Marginal probabilities: output bin b collects |psi[i]|^2 over the indices i
// whose bits at the measured positions spell b (first measured wire = most
// significant bit of b). One block sums one bin over a contiguous range of
// the complementary index space, walking it with the masked increment
// x' = ((x | ~mask) + step) & mask so consecutive threads read consecutive
// amplitudes whenever the low bits are unmeasured; a second kernel adds the
// ranges of each bin in a fixed order. With few amplitudes per bin a thread
// owns a bin instead.
//
// Sampling: the exclusive FP64 CDF of the full distribution by a three-phase
// scan (block totals, one-block scan of the totals, block-local scans), then
// one thread per shot: a SplitMix64 uniform u in (0, 1] (Steele, Lea, Flood
// 2014; output k of the generator seeded with `seed`) and the reference's
// binary search for the bin with cdf[k] < u <= cdf[k+1].
#include "device_common.hpp"

namespace plk {

namespace {

constexpr int PB = 256;          // threads per block
constexpr int MARG_R = 16;       // complement elements per thread
constexpr int SCAN_R = 16;       // contiguous elements per thread in the scan
constexpr int SCAN_CHUNK = PB * SCAN_R;

struct BinMap {
    int m;        // measured wires
    int g[64];    // global bit of measured wire k (k = 0 most significant)
};

__device__ __forceinline__ uint64_t bin_bits(const BinMap &bm, uint64_t b) {
    uint64_t v = 0;
    for (int k = 0; k < bm.m; k++) {
        v |= ((b >> (bm.m - 1 - k)) & 1ull) << bm.g[k];
    }
    return v;
}

__device__ __forceinline__ uint64_t pdep(uint64_t c, uint64_t mask) {
    uint64_t out = 0;
    while (mask) {
        const uint64_t low = mask & (~mask + 1ull);
        if (c & 1ull) {
            out |= low;
        }
        c >>= 1;
        mask &= mask - 1ull;
    }
    return out;
}

template <typename T> __device__ __forceinline__ double prob_of(const Cx<T> &a) {
    const double r = static_cast<double>(a.r), i = static_cast<double>(a.i);
    return r * r + i * i;
}

__device__ __forceinline__ double block_sum(double v, double *lds) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) {
        lds[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < PB / 64; w++) {
            t += lds[w];
        }
    }
    return t;
}

// One block per (bin, range of MARG_R * PB complement indices).
template <typename T>
__global__ void __launch_bounds__(PB)
    marg_block_kernel(const Cx<T> *__restrict__ psi, BinMap bm, uint64_t cmask, uint64_t step,
                      uint64_t count, uint64_t chunks, double *__restrict__ partial) {
    __shared__ double lds[PB / 64];
    const uint64_t b = blockIdx.x / chunks;
    const uint64_t chunk = blockIdx.x % chunks;
    const uint64_t base = bin_bits(bm, b);
    const uint64_t c0 = chunk * (uint64_t{MARG_R} * PB) + threadIdx.x;
    uint64_t x = pdep(c0, cmask);
    double acc = 0.0;
    for (int r = 0; r < MARG_R; r++) {
        if (c0 + static_cast<uint64_t>(r) * PB >= count) {
            break;
        }
        acc += prob_of(psi[base | x]);
        x = ((x | ~cmask) + step) & cmask;
    }
    const double t = block_sum(acc, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(PB)
    marg_finish_kernel(const double *__restrict__ partial, uint64_t bins, uint64_t chunks,
                       double *__restrict__ out) {
    const uint64_t b = uint64_t{blockIdx.x} * PB + threadIdx.x;
    if (b >= bins) {
        return;
    }
    double s = 0.0;
    for (uint64_t c = 0; c < chunks; c++) {
        s += partial[b * chunks + c];
    }
    out[b] = s;
}

// One thread per bin (few amplitudes per bin).
template <typename T>
__global__ void __launch_bounds__(PB)
    marg_thread_kernel(const Cx<T> *__restrict__ psi, BinMap bm, uint64_t cmask, uint64_t count,
                       uint64_t bins, double *__restrict__ out) {
    const uint64_t b = uint64_t{blockIdx.x} * PB + threadIdx.x;
    if (b >= bins) {
        return;
    }
    const uint64_t base = bin_bits(bm, b);
    uint64_t x = 0;
    double s = 0.0;
    for (uint64_t c = 0; c < count; c++) {
        s += prob_of(psi[base | x]);
        x = ((x | ~cmask) + 1ull) & cmask;
    }
    out[b] = s;
}

// Phase 1 of the CDF: total probability of each SCAN_CHUNK-element block.
template <typename T>
__global__ void __launch_bounds__(PB)
    scan_totals_kernel(const Cx<T> *__restrict__ psi, uint64_t len, double *__restrict__ totals) {
    __shared__ double lds[PB / 64];
    const uint64_t first = uint64_t{blockIdx.x} * SCAN_CHUNK + uint64_t{threadIdx.x} * SCAN_R;
    double s = 0.0;
    for (int r = 0; r < SCAN_R; r++) {
        if (first + r < len) {
            s += prob_of(psi[first + r]);
        }
    }
    const double t = block_sum(s, lds);
    if (threadIdx.x == 0) {
        totals[blockIdx.x] = t;
    }
}

// Phase 2: exclusive scan of the block totals by one block.
__global__ void __launch_bounds__(PB) scan_top_kernel(double *__restrict__ totals, uint64_t nb) {
    __shared__ double seg[PB];
    const uint64_t per = (nb + PB - 1) / PB;
    const uint64_t lo = threadIdx.x * per;
    const uint64_t hi = lo + per < nb ? lo + per : nb;
    double s = 0.0;
    for (uint64_t i = lo; i < hi; i++) {
        s += totals[i];
    }
    seg[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int t = 0; t < PB; t++) {
            const double v = seg[t];
            seg[t] = run;
            run += v;
        }
    }
    __syncthreads();
    double run = seg[threadIdx.x];
    for (uint64_t i = lo; i < hi; i++) {
        const double v = totals[i];
        totals[i] = run;
        run += v;
    }
}

// Phase 3: cdf[k] = sum_{i<k} |psi_i|^2.
template <typename T>
__global__ void __launch_bounds__(PB)
    scan_final_kernel(const Cx<T> *__restrict__ psi, uint64_t len,
                      const double *__restrict__ offsets, double *__restrict__ cdf) {
    __shared__ double seg[PB];
    const uint64_t first = uint64_t{blockIdx.x} * SCAN_CHUNK + uint64_t{threadIdx.x} * SCAN_R;
    double p[SCAN_R];
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < SCAN_R; r++) {
        p[r] = first + r < len ? prob_of(psi[first + r]) : 0.0;
        s += p[r];
    }
    seg[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = offsets[blockIdx.x];
        for (int t = 0; t < PB; t++) {
            const double v = seg[t];
            seg[t] = run;
            run += v;
        }
    }
    __syncthreads();
    double run = seg[threadIdx.x];
#pragma unroll
    for (int r = 0; r < SCAN_R; r++) {
        if (first + r < len) {
            cdf[first + r] = run;
        }
        run += p[r];
    }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// MeasuresFunctors.hpp:74-110: bin of u in the exclusive CDF, bits of the
// bin written most significant (wire 0) first.
__global__ void __launch_bounds__(PB)
    sample_kernel(const double *__restrict__ cdf, uint64_t len, int n, uint64_t shots,
                  uint64_t seed, uint64_t *__restrict__ out) {
    const uint64_t s = uint64_t{blockIdx.x} * PB + threadIdx.x;
    if (s >= shots) {
        return;
    }
    const uint64_t z = splitmix64(seed + (s + 1) * 0x9E3779B97F4A7C15ull);
    const double u = static_cast<double>((z >> 11) + 1) * 0x1.0p-53;
    uint64_t index;
    if (u <= cdf[1]) {
        index = 0;
    } else {
        uint64_t low = 1, high = len;
        while (high - low > 1) {
            const uint64_t mid = high - ((high - low) >> 1);
            const double c = mid == len ? 1.0 : cdf[mid];
            if (c < u) {
                low = mid;
            } else {
                high = mid;
            }
        }
        index = high - 1;
    }
    for (int j = 0; j < n; j++) {
        out[s * n + (n - 1 - j)] = (index >> j) & 1ull;
    }
}

template <typename T>
void probs_t(Device &dev, const State &s, const std::vector<int> &gbits, double *host_out) {
    const int m = static_cast<int>(gbits.size());
    BinMap bm{};
    bm.m = m;
    uint64_t wmask = 0;
    for (int k = 0; k < m; k++) {
        bm.g[k] = gbits[k];
        wmask |= uint64_t{1} << gbits[k];
    }
    const uint64_t full = s.n >= 64 ? ~uint64_t{0} : (uint64_t{1} << s.n) - 1;
    const uint64_t cmask = full & ~wmask;
    const uint64_t bins = uint64_t{1} << m;
    const uint64_t count = uint64_t{1} << (s.n - m);
    const auto *psi = static_cast<const Cx<T> *>(s.d);
    ProfScope prof("probs", static_cast<double>(s.bytes()), dev.stream);
    if (count <= 64) {
        double *out = static_cast<double *>(ensure_work(dev, bins * sizeof(double)));
        const unsigned grid = static_cast<unsigned>((bins + PB - 1) / PB);
        hipLaunchKernelGGL((marg_thread_kernel<T>), dim3(grid), dim3(PB), 0, dev.stream, psi, bm,
                           cmask, count, bins, out);
        PLK_HIP(hipGetLastError());
        PLK_HIP(hipMemcpyAsync(host_out, out, bins * sizeof(double), hipMemcpyDeviceToHost,
                               dev.stream));
    } else {
        const uint64_t per_block = uint64_t{MARG_R} * PB;
        const uint64_t chunks = (count + per_block - 1) / per_block;
        if (bins * chunks > (uint64_t{1} << 31)) {
            raise(PLK_ERR_UNSUPPORTED, "probability grid too large");
        }
        double *partial = static_cast<double *>(
            ensure_work(dev, (bins * chunks + bins) * sizeof(double)));
        double *out = partial + bins * chunks;
        const uint64_t step = [&] {
            uint64_t v = 0, c = PB, mk = cmask;
            while (mk) {
                const uint64_t low = mk & (~mk + 1);
                if (c & 1) {
                    v |= low;
                }
                c >>= 1;
                mk &= mk - 1;
            }
            return v;
        }();
        hipLaunchKernelGGL((marg_block_kernel<T>), dim3(static_cast<unsigned>(bins * chunks)),
                           dim3(PB), 0, dev.stream, psi, bm, cmask, step, count, chunks, partial);
        PLK_HIP(hipGetLastError());
        hipLaunchKernelGGL(marg_finish_kernel, dim3(static_cast<unsigned>((bins + PB - 1) / PB)),
                           dim3(PB), 0, dev.stream, partial, bins, chunks, out);
        PLK_HIP(hipGetLastError());
        PLK_HIP(hipMemcpyAsync(host_out, out, bins * sizeof(double), hipMemcpyDeviceToHost,
                               dev.stream));
    }
    PLK_HIP(hipStreamSynchronize(dev.stream));
}

template <typename T>
void sample_t(Device &dev, const State &s, size_t shots, uint64_t seed, size_t *host_out) {
    const uint64_t len = s.len;
    const uint64_t nb = (len + SCAN_CHUNK - 1) / SCAN_CHUNK;
    const size_t cdf_bytes = len * sizeof(double);
    const size_t tot_bytes = ((nb * sizeof(double)) + 255) & ~size_t{255};
    const size_t out_bytes = static_cast<size_t>(shots) * s.n * sizeof(uint64_t);
    unsigned char *w = static_cast<unsigned char *>(
        ensure_work(dev, ((cdf_bytes + 255) & ~size_t{255}) + tot_bytes + out_bytes));
    double *cdf = reinterpret_cast<double *>(w);
    double *totals = reinterpret_cast<double *>(w + ((cdf_bytes + 255) & ~size_t{255}));
    uint64_t *out = reinterpret_cast<uint64_t *>(reinterpret_cast<unsigned char *>(totals) +
                                                 tot_bytes);
    const auto *psi = static_cast<const Cx<T> *>(s.d);
    ProfScope prof("sample", 3.0 * static_cast<double>(s.bytes()), dev.stream);
    hipLaunchKernelGGL((scan_totals_kernel<T>), dim3(static_cast<unsigned>(nb)), dim3(PB), 0,
                       dev.stream, psi, len, totals);
    PLK_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(PB), 0, dev.stream, totals, nb);
    PLK_HIP(hipGetLastError());
    hipLaunchKernelGGL((scan_final_kernel<T>), dim3(static_cast<unsigned>(nb)), dim3(PB), 0,
                       dev.stream, psi, len, totals, cdf);
    PLK_HIP(hipGetLastError());
    if (shots > 0) {
        hipLaunchKernelGGL(sample_kernel, dim3(static_cast<unsigned>((shots + PB - 1) / PB)),
                           dim3(PB), 0, dev.stream, cdf, len, s.n, static_cast<uint64_t>(shots),
                           seed, out);
        PLK_HIP(hipGetLastError());
        PLK_HIP(hipMemcpyAsync(host_out, out, out_bytes, hipMemcpyDeviceToHost, dev.stream));
    }
    PLK_HIP(hipStreamSynchronize(dev.stream));
}

} // namespace

void probs(Device &dev, const State &s, const std::vector<int> &gbits, double *host_out) {
    if (s.dtype == PLK_C64) {
        probs_t<float>(dev, s, gbits, host_out);
    } else {
        probs_t<double>(dev, s, gbits, host_out);
    }
}

void sample(Device &dev, const State &s, size_t shots, uint64_t seed, size_t *host_out) {
    if (s.dtype == PLK_C64) {
        sample_t<float>(dev, s, shots, seed, host_out);
    } else {
        sample_t<double>(dev, s, shots, seed, host_out);
    }
}

} // namespace plk
