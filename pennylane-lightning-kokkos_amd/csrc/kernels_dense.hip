// Dense matrices on up to 12 wires: <psi|M|psi> and out = M psi.
//
// The reference evaluates a Hamiltonian on fewer than 13 wires as one dense
// matrix (lightning_kokkos.py:521-534 -> MeasuresKokkos.hpp:283-346, which
// copies the state and applies the matrix through multiQubitOpFunctor,
// GateFunctors.hpp:195-300). Here one block takes one group of amplitudes
// that share the bits outside the m wires (2^m amplitudes, staged in LDS,
// at most 64 KiB) and a chunk of 256 output rows; thread r forms
// w_r = sum_c M[r][c] v_c reading the transposed matrix so that consecutive
// threads read consecutive entries. The expectation value accumulates
// Re(conj(v_r) w_r) in FP64 (fixed-order block sums); the application writes
// w_r to a separate output state.
//
// Bytes: 4^m matrix entries per group (the matrix is re-read by every group
// and row chunk, from L2 when it fits) plus 2^n amplitudes read (and written
// for the application).
#include "plk_internal.hpp"

namespace plk {

namespace {

template <typename T> struct CxD {
    T r, i;
};

constexpr int kDenseBlock = 256;

struct DenseBigArgs {
    int m;           // wires
    int bits[12];    // ascending global bits of the wires
    int pad;
    uint64_t groups; // 2^(n - m)
};

// Local index k of the matrix (wires[0] most significant) to global offset.
__device__ __forceinline__ uint64_t group_base(const DenseBigArgs &a, uint64_t g) {
    for (int q = 0; q < a.m; q++) {
        const int b = a.bits[q];
        const uint64_t lo = g & ((uint64_t{1} << b) - 1);
        g = ((g >> b) << (b + 1)) | lo;
    }
    return g;
}

template <typename T, bool EXPVAL>
__global__ void __launch_bounds__(kDenseBlock)
    dense_big_kernel(const CxD<T> *__restrict__ psi, CxD<T> *__restrict__ out,
                     const CxD<T> *__restrict__ mt, const uint64_t *__restrict__ off,
                     DenseBigArgs a, double *__restrict__ partials) {
    extern __shared__ unsigned char smem[];
    auto *v = reinterpret_cast<CxD<T> *>(smem);
    const int dim = 1 << a.m;
    auto *red = reinterpret_cast<double *>(smem + static_cast<size_t>(dim) * sizeof(CxD<T>));
    const int r = blockIdx.x * kDenseBlock + threadIdx.x;
    double acc = 0.0;
    for (uint64_t g = blockIdx.y; g < a.groups; g += gridDim.y) {
        const uint64_t base = group_base(a, g);
        __syncthreads(); // the previous group's rows are done with v
        for (int k = threadIdx.x; k < dim; k += kDenseBlock) {
            v[k] = psi[base + off[k]];
        }
        __syncthreads();
        if (r < dim) {
            double wr = 0.0, wi = 0.0;
            for (int c = 0; c < dim; c++) {
                const CxD<T> m = mt[static_cast<uint64_t>(c) * dim + r];
                const CxD<T> x = v[c];
                wr = fma(static_cast<double>(m.r), static_cast<double>(x.r),
                         fma(-static_cast<double>(m.i), static_cast<double>(x.i), wr));
                wi = fma(static_cast<double>(m.r), static_cast<double>(x.i),
                         fma(static_cast<double>(m.i), static_cast<double>(x.r), wi));
            }
            if (EXPVAL) {
                acc += static_cast<double>(v[r].r) * wr + static_cast<double>(v[r].i) * wi;
            } else {
                out[base + off[r]] = CxD<T>{static_cast<T>(wr), static_cast<T>(wi)};
            }
        }
    }
    if (EXPVAL) {
        // Fixed-order block sum: wave shuffles, then the four waves in order.
        for (int o = 32; o > 0; o >>= 1) {
            acc += __shfl_xor(acc, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            red[threadIdx.x >> 6] = acc;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int w = 0; w < kDenseBlock / 64; w++) {
                s += red[w];
            }
            partials[uint64_t{blockIdx.y} * gridDim.x + blockIdx.x] = s;
        }
    }
}

__global__ void __launch_bounds__(kDenseBlock)
    dense_sum_kernel(const double *__restrict__ partials, uint64_t count, double *__restrict__ out) {
    double acc = 0.0;
    for (uint64_t i = threadIdx.x; i < count; i += kDenseBlock) {
        acc += partials[i];
    }
    __shared__ double red[kDenseBlock / 64];
    for (int o = 32; o > 0; o >>= 1) {
        acc += __shfl_xor(acc, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kDenseBlock / 64; w++) {
            s += red[w];
        }
        out[0] = s;
    }
}

// Lets the kernels take more than 64 KiB of LDS (a 12-wire complex128 group
// is exactly 64 KiB before the reduction slots).
template <typename T> void set_dense_lds_limit() {
    static bool done = false;
    if (!done) {
        const int bytes = static_cast<int>((size_t{1} << kDenseMaxWires) * sizeof(CxD<T>) + 1024);
        PLK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&dense_big_kernel<T, true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        PLK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&dense_big_kernel<T, false>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        done = true;
    }
}

template <typename T>
double dense_big_t(Device &dev, const State &s, const State *out, const std::vector<size_t> &wires,
                   const std::vector<cd> &mat) {
    const int m = static_cast<int>(wires.size());
    const size_t dim = size_t{1} << m;
    DenseBigArgs a{};
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
    a.groups = uint64_t{1} << (s.n - m);
    // Transposed matrix and local-index offsets in one device block.
    const size_t mt_bytes = dim * dim * sizeof(CxD<T>);
    const size_t off_at = (mt_bytes + 255) / 256 * 256;
    const size_t part_at = off_at + ((dim * sizeof(uint64_t) + 255) / 256 * 256);
    const uint64_t row_chunks = (dim + kDenseBlock - 1) / kDenseBlock;
    const uint64_t gy = std::min<uint64_t>(a.groups, static_cast<uint64_t>(dev.num_cus) * 8);
    const size_t nparts = static_cast<size_t>(row_chunks * gy);
    const size_t total = part_at + (nparts + 1) * sizeof(double);
    std::vector<unsigned char> host(part_at);
    auto *mt = reinterpret_cast<CxD<T> *>(host.data());
    for (size_t r = 0; r < dim; r++) {
        for (size_t c = 0; c < dim; c++) {
            const cd v = mat[r * dim + c];
            mt[c * dim + r] = CxD<T>{static_cast<T>(v.real()), static_cast<T>(v.imag())};
        }
    }
    auto *off = reinterpret_cast<uint64_t *>(host.data() + off_at);
    for (size_t k = 0; k < dim; k++) {
        uint64_t o = 0;
        for (int p = 0; p < m; p++) {
            if ((k >> (m - 1 - p)) & 1U) {
                o |= uint64_t{1} << bits[p];
            }
        }
        off[k] = o;
    }
    auto *d = static_cast<unsigned char *>(arena_alloc(dev, total));
    PLK_HIP(hipMemcpyAsync(d, host.data(), host.size(), hipMemcpyHostToDevice, dev.stream));
    const dim3 grid(static_cast<unsigned>(row_chunks), static_cast<unsigned>(gy));
    const size_t lds = dim * sizeof(CxD<T>) + (kDenseBlock / 64) * sizeof(double);
    set_dense_lds_limit<T>();
    auto *parts = reinterpret_cast<double *>(d + part_at);
    double result = 0.0;
    if (out == nullptr) {
        ProfScope prof("dense_expval", static_cast<double>(a.groups) * mt_bytes + s.bytes(),
                       dev.stream);
        hipLaunchKernelGGL((dense_big_kernel<T, true>), grid, dim3(kDenseBlock), lds, dev.stream,
                           static_cast<const CxD<T> *>(s.d), nullptr,
                           reinterpret_cast<const CxD<T> *>(d), reinterpret_cast<const uint64_t *>(d + off_at),
                           a, parts);
        PLK_HIP(hipGetLastError());
        hipLaunchKernelGGL(dense_sum_kernel, dim3(1), dim3(kDenseBlock), 0, dev.stream, parts,
                           static_cast<uint64_t>(nparts), parts + nparts);
        PLK_HIP(hipGetLastError());
        PLK_HIP(hipMemcpyAsync(&result, parts + nparts, sizeof(double), hipMemcpyDeviceToHost,
                               dev.stream));
    } else {
        ProfScope prof("dense_apply", static_cast<double>(a.groups) * mt_bytes + 2.0 * s.bytes(),
                       dev.stream);
        hipLaunchKernelGGL((dense_big_kernel<T, false>), grid, dim3(kDenseBlock), lds, dev.stream,
                           static_cast<const CxD<T> *>(s.d), static_cast<CxD<T> *>(out->d),
                           reinterpret_cast<const CxD<T> *>(d), reinterpret_cast<const uint64_t *>(d + off_at),
                           a, parts);
        PLK_HIP(hipGetLastError());
    }
    // The host image is read by the copy above.
    PLK_HIP(hipStreamSynchronize(dev.stream));
    return result;
}

void check_dense_big(const State &s, const std::vector<size_t> &wires, const std::vector<cd> &mat) {
    const size_t m = wires.size();
    if (m == 0 || m > kDenseMaxWires || static_cast<int>(m) > s.n) {
        raise(PLK_ERR_UNSUPPORTED, "dense matrices act on 1.." + std::to_string(kDenseMaxWires) +
                                       " wires of the state");
    }
    if (mat.size() != (size_t{1} << (2 * m))) {
        raise(PLK_ERR_INVALID, "dense matrix size does not match its wires");
    }

}

} // namespace

double dense_expval(Device &dev, const State &s, const std::vector<size_t> &wires,
                    const std::vector<cd> &mat) {
    check_dense_big(s, wires, mat);
    return s.dtype == PLK_C64 ? dense_big_t<float>(dev, s, nullptr, wires, mat)
                              : dense_big_t<double>(dev, s, nullptr, wires, mat);
}

void dense_apply_into(Device &dev, const State &out, const State &in,
                      const std::vector<size_t> &wires, const std::vector<cd> &mat) {
    check_dense_big(in, wires, mat);
    if (out.len != in.len || out.dtype != in.dtype) {
        raise(PLK_ERR_INVALID, "dense application between different shapes");
    }
    if (in.dtype == PLK_C64) {
        dense_big_t<float>(dev, in, &out, wires, mat);
    } else {
        dense_big_t<double>(dev, in, &out, wires, mat);
    }
}

} // namespace plk
