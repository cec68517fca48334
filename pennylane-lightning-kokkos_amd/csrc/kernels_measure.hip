// Measurement kernels: FP64 inner products, Pauli-word sums (expectation
// value and application) and CSR sparse matrices.
//
// Pauli words: for P = (x)_w sigma_w with X-mask x (X or Y on the wire),
// sign mask zy (Z or Y on the wire) and n_Y Y factors,
//   (P psi)[i] = (-i)^{n_Y} (-1)^{popcount(i & zy)} psi[i ^ x].
// Terms sharing an X mask share the gathered amplitude psi[i ^ x], so a
// Hamiltonian is evaluated with one gather per distinct X mask instead of the
// reference's per-term state copy, gate kernels and axpy
// (simulator/ObservablesKokkos.hpp:360-373; MeasuresKokkos.hpp:354-360).
#include "plk_internal.hpp"

namespace plk {

namespace {

template <typename T> struct Cx {
    T r, i;
};
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
__device__ __forceinline__ void st(Cx<T> *p, uint64_t i, Cx<T> c) {
    using V = typename VecOf<T>::type;
    V v;
    v.x = c.r;
    v.y = c.i;
    reinterpret_cast<V *>(p)[i] = v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, 64);
    }
    return v;
}

// Block-wide FP64 sum of NV values; thread 0 receives the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double red[4][NV > 0 ? NV : 1];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double s = wave_sum(v[k]);
        if (lane == 0) {
            red[wave][k] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            v[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
        }
    }
}

constexpr int kBlock = 256;

// sum_i conj(x_j[i]) y[i] for j < NX; partials[(2*j + {0,1}) * grid + block].
template <typename T, int NX> struct XPtrs {
    const Cx<T> *p[NX];
};

template <typename T, int NX>
__global__ void __launch_bounds__(kBlock)
    dot_kernel(XPtrs<T, NX> xs, const Cx<T> *__restrict__ y, uint64_t len,
               double *__restrict__ partials) {
    double acc[2 * NX];
#pragma unroll
    for (int k = 0; k < 2 * NX; k++) {
        acc[k] = 0.0;
    }
    for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < len;
         i += uint64_t{gridDim.x} * blockDim.x) {
        const Cx<T> b = ld(y, i);
        const double br = b.r, bi = b.i;
#pragma unroll
        for (int j = 0; j < NX; j++) {
            const Cx<T> a = ld(xs.p[j], i);
            const double ar = a.r, ai = a.i;
            acc[2 * j] = fma(ar, br, fma(ai, bi, acc[2 * j]));
            acc[2 * j + 1] = fma(ar, bi, fma(-ai, br, acc[2 * j + 1]));
        }
    }
    block_sum<2 * NX>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 2 * NX; k++) {
            partials[uint64_t{static_cast<uint32_t>(k)} * gridDim.x + blockIdx.x] = acc[k];
        }
    }
}

// out[k] = sum_b partials[k * grid + b] for k < nvals (one block per value).
__global__ void __launch_bounds__(kBlock)
    final_sum_kernel(const double *__restrict__ partials, int grid,
                     double *__restrict__ out) {
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < grid; b += blockDim.x) {
        v[0] += partials[uint64_t{blockIdx.x} * grid + b];
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) {
        out[blockIdx.x] = v[0];
    }
}

struct PauliGroup {
    uint64_t x;
    int first;
    int count;
};
struct PauliTermDev {
    uint64_t zy;
    double cr, ci;
};

// Coefficient of psi[i ^ x] in (H psi)[i] for one X-mask group.
__device__ __forceinline__ void group_coeff(const PauliTermDev *__restrict__ terms,
                                            const PauliGroup &g, uint64_t i, double &cr,
                                            double &ci) {
    cr = 0.0;
    ci = 0.0;
    for (int k = 0; k < g.count; k++) {
        const PauliTermDev t = terms[g.first + k];
        const double sgn = (__popcll(i & t.zy) & 1) ? -1.0 : 1.0;
        cr = fma(sgn, t.cr, cr);
        ci = fma(sgn, t.ci, ci);
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    pauli_expval_kernel(const Cx<T> *__restrict__ psi, uint64_t len,
                        const PauliGroup *__restrict__ groups, int ngroups,
                        const PauliTermDev *__restrict__ terms,
                        double *__restrict__ partials) {
    // H is Hermitian, so for an X mask x != 0 the terms at i and at i ^ x
    // contribute equally: only indices whose highest bit of x is 0 are
    // visited, counted twice (whole waves skip together when that bit is
    // above the wave's index bits).
    double acc[1] = {0.0};
    for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < len;
         i += uint64_t{gridDim.x} * blockDim.x) {
        const Cx<T> a = ld(psi, i);
        const double ar = a.r, ai = a.i;
        for (int gi = 0; gi < ngroups; gi++) {
            const PauliGroup g = groups[gi];
            double w = 1.0;
            if (g.x) {
                const uint64_t top = uint64_t{1} << (63 - __clzll(static_cast<long long>(g.x)));
                if (i & top) {
                    continue;
                }
                w = 2.0;
            }
            const Cx<T> b = g.x ? ld(psi, i ^ g.x) : a;
            double cr, ci;
            group_coeff(terms, g, i, cr, ci);
            // Re(conj(a) * c * b)
            const double br = b.r, bi = b.i;
            const double tr = cr * br - ci * bi;
            const double ti = cr * bi + ci * br;
            acc[0] = fma(w * ar, tr, fma(w * ai, ti, acc[0]));
        }
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = acc[0];
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    pauli_apply_kernel(Cx<T> *__restrict__ out, const Cx<T> *__restrict__ psi,
                       uint64_t len, const PauliGroup *__restrict__ groups, int ngroups,
                       const PauliTermDev *__restrict__ terms) {
    for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < len;
         i += uint64_t{gridDim.x} * blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int gi = 0; gi < ngroups; gi++) {
            const PauliGroup g = groups[gi];
            const Cx<T> b = ld(psi, i ^ g.x);
            double cr, ci;
            group_coeff(terms, g, i, cr, ci);
            const double br = b.r, bi = b.i;
            sr += cr * br - ci * bi;
            si += cr * bi + ci * br;
        }
        st(out, i, Cx<T>{static_cast<T>(sr), static_cast<T>(si)});
    }
}

// CSR kernels: one wave works on R = 64 / G consecutive rows at a time, G
// lanes per row striding over the row's entries, so a wave's loads of the
// values and indices are contiguous runs (streamed with non-temporal loads:
// they are read once), while the gathers psi[col] stay cacheable. Each block
// owns one contiguous range of `chunk` rows and its four waves take
// consecutive row groups of it: rows that gather from the same cache lines
// (a Pauli-sum matrix has col = row ^ x, so 8 consecutive complex128 rows
// share every line) run on one CU at about the same time instead of being
// spread over the XCDs' separate L2s. Every row range and column index is
// checked: a malformed matrix is skipped and counted in the block's flag slot
// instead of being read out of bounds.
template <typename V> struct Vec2;
template <> struct Vec2<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template <> struct Vec2<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};

template <typename V> __device__ __forceinline__ Cx<V> ld_stream(const Cx<V> *p, uint64_t k) {
    using W = typename Vec2<V>::type;
    const W w = __builtin_nontemporal_load(reinterpret_cast<const W *>(p) + k);
    return {w.x, w.y};
}

template <typename I> __device__ __forceinline__ int64_t ld_index(const I *p, uint64_t k) {
    return static_cast<int64_t>(__builtin_nontemporal_load(p + k));
}

// Partial sum over the row entries start + sub, start + sub + G, ... of
// A_rk psi[c_k]. Four entries per lane are loaded before any is used, so a
// lane has four index loads and then four gathers in flight instead of one
// dependent pair at a time. Out-of-range columns set bad.
template <typename T, typename V, typename I, int G>
__device__ __forceinline__ void csr_row_sum(const Cx<T> *__restrict__ psi, uint64_t len,
                                            const Cx<V> *__restrict__ data,
                                            const I *__restrict__ indices, int64_t start,
                                            int64_t end, int sub, double &sr, double &si,
                                            double &bad) {
    constexpr int U = 4;
    for (int64_t k0 = start + sub; k0 < end; k0 += U * G) {
        int64_t c[U];
        Cx<V> a[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t k = k0 + static_cast<int64_t>(u) * G;
            c[u] = k < end ? ld_index(indices, static_cast<uint64_t>(k)) : -1;
            a[u] = k < end ? ld_stream(data, static_cast<uint64_t>(k)) : Cx<V>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t k = k0 + static_cast<int64_t>(u) * G;
            if (k >= end) {
                continue;
            }
            if (c[u] < 0 || static_cast<uint64_t>(c[u]) >= len) {
                bad = 1.0;
                continue;
            }
            const Cx<T> b = ld(psi, static_cast<uint64_t>(c[u]));
            const double ar = a[u].r, ai = a[u].i, br = b.r, bi = b.i;
            sr = fma(ar, br, fma(-ai, bi, sr));
            si = fma(ar, bi, fma(ai, br, si));
        }
    }
}

template <typename T, typename V, typename I, int G>
__global__ void __launch_bounds__(kBlock)
    csr_expval_kernel(const Cx<T> *__restrict__ psi, uint64_t len,
                      const Cx<V> *__restrict__ data, const I *__restrict__ indices,
                      const I *__restrict__ indptr, uint64_t rows, uint64_t nnz,
                      uint64_t chunk, double *__restrict__ partials) {
    constexpr int R = 64 / G;
    constexpr int W = kBlock / 64;
    const int lane = threadIdx.x & 63;
    const int sub = lane & (G - 1);
    const uint64_t first = uint64_t{blockIdx.x} * chunk;
    const uint64_t last = first + chunk < rows ? first + chunk : rows;
    double acc[2] = {0.0, 0.0};
    for (uint64_t r = first + (threadIdx.x >> 6) * R + lane / G; r < last; r += W * R) {
        const int64_t start = indptr[r], end = indptr[r + 1];
        if (start < 0 || end < start || static_cast<uint64_t>(end) > nnz) {
            acc[1] = 1.0;
            continue;
        }
        double sr = 0.0, si = 0.0;
        csr_row_sum<T, V, I, G>(psi, len, data, indices, start, end, sub, sr, si, acc[1]);
        // Re(conj(psi_r) * partial row sum): summed over the row's lanes later.
        const Cx<T> p = ld(psi, r);
        acc[0] = fma(static_cast<double>(p.r), sr, fma(static_cast<double>(p.i), si, acc[0]));
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = acc[0];
        partials[gridDim.x + blockIdx.x] = acc[1];
    }
}

template <typename T, typename V, typename I, int G>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_kernel(Cx<T> *__restrict__ out, const Cx<T> *__restrict__ psi, uint64_t len,
                    const Cx<V> *__restrict__ data, const I *__restrict__ indices,
                    const I *__restrict__ indptr, uint64_t rows, uint64_t nnz,
                    uint64_t chunk, double *__restrict__ partials) {
    constexpr int R = 64 / G;
    constexpr int W = kBlock / 64;
    const int lane = threadIdx.x & 63;
    const int sub = lane & (G - 1);
    const uint64_t first = uint64_t{blockIdx.x} * chunk;
    const uint64_t last = first + chunk < rows ? first + chunk : rows;
    double bad[1] = {0.0};
    // Wave-uniform loop: every lane takes part in the row-sum shuffles.
    for (uint64_t r0 = first + (threadIdx.x >> 6) * R; r0 < last; r0 += W * R) {
        const uint64_t r = r0 + lane / G;
        double sr = 0.0, si = 0.0;
        if (r < last) {
            const int64_t start = indptr[r], end = indptr[r + 1];
            if (start < 0 || end < start || static_cast<uint64_t>(end) > nnz) {
                bad[0] = 1.0;
            } else {
                csr_row_sum<T, V, I, G>(psi, len, data, indices, start, end, sub, sr, si, bad[0]);
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off, 64);
            si += __shfl_xor(si, off, 64);
        }
        if (sub == 0 && r < last) {
            st(out, r, Cx<T>{static_cast<T>(sr), static_cast<T>(si)});
        }
    }
    block_sum<1>(bad);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = bad[0];
    }
}

unsigned grid_for(const Device &dev, uint64_t work) {
    const uint64_t want = (work + kBlock - 1) / kBlock;
    const uint64_t cap = static_cast<uint64_t>(dev.num_cus) * 4;
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min(want, cap)));
}

// Sums nvals partial rows of length grid and copies the totals to the host.
void finish_sums(Device &dev, int grid, int nvals, double *host_out) {
    hipLaunchKernelGGL(final_sum_kernel, dim3(nvals), dim3(kBlock), 0, dev.stream, dev.red,
                       grid, dev.red_out);
    PLK_HIP(hipGetLastError());
    PLK_HIP(hipMemcpyAsync(host_out, dev.red_out, sizeof(double) * nvals,
                           hipMemcpyDeviceToHost, dev.stream));
    PLK_HIP(hipStreamSynchronize(dev.stream));
}

template <typename T, int NX>
void dot_chunk(Device &dev, const State *const *xs, const State &y, double *out) {
    XPtrs<T, NX> p{};
    for (int j = 0; j < NX; j++) {
        p.p[j] = static_cast<const Cx<T> *>(xs[j]->d);
    }
    const unsigned grid = grid_for(dev, y.len);
    ensure_red(dev, static_cast<size_t>(grid) * 2 * NX, 2 * NX);
    {
        ProfScope prof("inner_product", static_cast<double>((NX + 1) * y.bytes()), dev.stream);
        hipLaunchKernelGGL((dot_kernel<T, NX>), dim3(grid), dim3(kBlock), 0, dev.stream, p,
                           static_cast<const Cx<T> *>(y.d), static_cast<uint64_t>(y.len),
                           dev.red);
        PLK_HIP(hipGetLastError());
    }
    finish_sums(dev, static_cast<int>(grid), 2 * NX, out);
}

template <typename T>
void inner_products_t(Device &dev, const State *const *xs, int nx, const State &y,
                      double *out) {
    int j = 0;
    while (j < nx) {
        const int left = nx - j;
        if (left >= 8) {
            dot_chunk<T, 8>(dev, xs + j, y, out + 2 * j);
            j += 8;
        } else if (left >= 4) {
            dot_chunk<T, 4>(dev, xs + j, y, out + 2 * j);
            j += 4;
        } else if (left >= 2) {
            dot_chunk<T, 2>(dev, xs + j, y, out + 2 * j);
            j += 2;
        } else {
            dot_chunk<T, 1>(dev, xs + j, y, out + 2 * j);
            j += 1;
        }
    }
}

// Uploads the grouped term tables; returns device pointers inside one block.
struct PauliTables {
    void *blob = nullptr;
    const PauliGroup *groups = nullptr;
    const PauliTermDev *terms = nullptr;
    int ngroups = 0;
};

// Host image of the grouped term tables: groups sorted by X mask, then the
// terms of each group, at byte offset toff.
struct PauliBlob {
    std::vector<unsigned char> bytes;
    size_t toff = 0;
    int ngroups = 0;
};

PauliBlob pack_pauli(const std::vector<PauliTerm> &terms) {
    std::vector<PauliTerm> sorted = terms;
    std::stable_sort(sorted.begin(), sorted.end(),
                     [](const PauliTerm &a, const PauliTerm &b) { return a.x < b.x; });
    std::vector<PauliGroup> groups;
    std::vector<PauliTermDev> tdev;
    for (size_t k = 0; k < sorted.size(); k++) {
        if (groups.empty() || groups.back().x != sorted[k].x) {
            groups.push_back({sorted[k].x, static_cast<int>(k), 0});
        }
        groups.back().count++;
        tdev.push_back({sorted[k].zy, sorted[k].coeff.real(), sorted[k].coeff.imag()});
    }
    PauliBlob b;
    const size_t gb = groups.size() * sizeof(PauliGroup);
    b.toff = (gb + 255) / 256 * 256;
    b.bytes.resize(b.toff + tdev.size() * sizeof(PauliTermDev) + 16);
    std::memcpy(b.bytes.data(), groups.data(), gb);
    std::memcpy(b.bytes.data() + b.toff, tdev.data(), tdev.size() * sizeof(PauliTermDev));
    b.ngroups = static_cast<int>(groups.size());
    return b;
}

PauliTables tables_at(void *dst, const PauliBlob &b) {
    PauliTables t;
    t.blob = dst;
    t.groups = static_cast<const PauliGroup *>(dst);
    t.terms = reinterpret_cast<const PauliTermDev *>(static_cast<unsigned char *>(dst) + b.toff);
    t.ngroups = b.ngroups;
    return t;
}

// Launches pauli_expval_kernel over grid blocks into partials.
void launch_pauli_expval(Device &dev, const State &s, const PauliTables &t, unsigned grid,
                         double *partials) {
    ProfScope prof("pauli_expval", static_cast<double>(t.ngroups) * s.bytes(), dev.stream);
    if (s.dtype == PLK_C64) {
        hipLaunchKernelGGL((pauli_expval_kernel<float>), dim3(grid), dim3(kBlock), 0, dev.stream,
                           static_cast<const Cx<float> *>(s.d), static_cast<uint64_t>(s.len),
                           t.groups, t.ngroups, t.terms, partials);
    } else {
        hipLaunchKernelGGL((pauli_expval_kernel<double>), dim3(grid), dim3(kBlock), 0,
                           dev.stream, static_cast<const Cx<double> *>(s.d),
                           static_cast<uint64_t>(s.len), t.groups, t.ngroups, t.terms, partials);
    }
    PLK_HIP(hipGetLastError());
}


} // namespace

void inner_products(Device &dev, const State *const *xs, int nx, const State &y,
                    double *out_re_im) {
    if (y.dtype == PLK_C64) {
        inner_products_t<float>(dev, xs, nx, y, out_re_im);
    } else {
        inner_products_t<double>(dev, xs, nx, y, out_re_im);
    }
}

PauliDevice::~PauliDevice() {
    if (blob != nullptr) {
        (void)hipSetDevice(device);
        (void)hipStreamSynchronize(stream);
        (void)hipFree(blob);
    }
}

std::unique_ptr<PauliDevice> pauli_upload(Device &dev, const std::vector<PauliTerm> &terms) {
    auto t = std::make_unique<PauliDevice>();
    t->device = dev.id;
    t->stream = dev.stream;
    if (terms.empty()) {
        return t;
    }
    const PauliBlob b = pack_pauli(terms);
    t->blob = dev_malloc(dev.id, b.bytes.size(), "Pauli tables");
    PLK_HIP(hipMemcpy(t->blob, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice));
    t->toff = b.toff;
    t->ngroups = b.ngroups;
    return t;
}

namespace {

PauliTables view_at(const void *blob, size_t toff, int ngroups) {
    PauliTables v;
    v.blob = const_cast<void *>(blob);
    v.groups = static_cast<const PauliGroup *>(blob);
    v.terms = reinterpret_cast<const PauliTermDev *>(static_cast<const unsigned char *>(blob) +
                                                      toff);
    v.ngroups = ngroups;
    return v;
}

PauliTables view_of(const PauliDevice &t) { return view_at(t.blob, t.toff, t.ngroups); }

} // namespace

double pauli_expval(Device &dev, const State &s, const PauliDevice &t) {
    if (t.ngroups == 0) {
        return 0.0;
    }
    const unsigned grid = grid_for(dev, s.len);
    ensure_red(dev, grid, 1);
    launch_pauli_expval(dev, s, view_of(t), grid, dev.red);
    double out = 0.0;
    finish_sums(dev, static_cast<int>(grid), 1, &out);
    return out;
}

double pauli_expval(Device &dev, const State &s, const std::vector<PauliTerm> &terms) {
    if (terms.empty()) {
        return 0.0;
    }
    return pauli_expval(dev, s, *pauli_upload(dev, terms));
}

// ---------------------------------------------------------------------------
// Allocation-free Pauli-sum expectation value.
//
// Every device this library runs on holds one copy of `scratch` until the
// process ends. Layout: the per-block partial sums (kNoAllocPartials
// doubles), the final sum (one double, own 256-byte line), then the grouped
// term tables. Host code never passes the array's name to a kernel: the name
// is a host-side stand-in, and the device address of each GPU's copy is
// looked up once per device with hipGetSymbolAddress and kept in
// Device::noalloc_scratch. Each call overwrites all of the parts it reads
// (every block writes its partial, final_sum_kernel writes the total, the
// tables are copied in), so nothing has to start at zero. All calls on one
// device share the array: two calls that overlap on a device race. The
// library's entry-point lock serialises its own calls and each call waits for
// its kernels before returning.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocPartials = 2048;
constexpr size_t kNoAllocTableOff = kNoAllocPartials * sizeof(double) + 256;
constexpr size_t kNoAllocScratchBytes = size_t{256} << 10;
static_assert(kNoAllocTableOff + 256 + kNoAllocMaxPauliWords * (sizeof(PauliGroup) + sizeof(PauliTermDev)) + 16 <=
                  kNoAllocScratchBytes,
              "scratch too small for the largest accepted table");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static unsigned char *scratch_of(Device &dev) {
    if (dev.noalloc_scratch == nullptr) {
        // device() made dev current.
        void *ptr = nullptr;
        PLK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr) {
            raise(PLK_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        }
        dev.noalloc_scratch = ptr;
    }
    return static_cast<unsigned char *>(dev.noalloc_scratch);
}

double pauli_expval_noalloc(Device &dev, const State &s, const std::vector<PauliTerm> &terms) {
    if (terms.size() > kNoAllocMaxPauliWords) {
        raise(PLK_ERR_INVALID, "plk_expval_obs_noalloc: the observable expands to " +
                                   std::to_string(terms.size()) + " Pauli words; at most " +
                                   std::to_string(kNoAllocMaxPauliWords) +
                                   " fit the library's fixed scratch array");
    }
    if (terms.empty()) {
        return 0.0;
    }
    unsigned char *base = scratch_of(dev);
    double *partials = reinterpret_cast<double *>(base);
    double *total = reinterpret_cast<double *>(base + kNoAllocPartials * sizeof(double));
    const PauliBlob b = pack_pauli(terms);
    if (kNoAllocTableOff + b.bytes.size() > kNoAllocScratchBytes) {
        raise(PLK_ERR_INVALID, "plk_expval_obs_noalloc: term tables exceed the scratch array");
    }
    PLK_HIP(hipMemcpyAsync(base + kNoAllocTableOff, b.bytes.data(), b.bytes.size(),
                           hipMemcpyHostToDevice, dev.stream));
    const PauliTables t = tables_at(base + kNoAllocTableOff, b);
    const unsigned grid = static_cast<unsigned>(
        std::min<uint64_t>(grid_for(dev, s.len), kNoAllocPartials));
    launch_pauli_expval(dev, s, t, grid, partials);
    hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(kBlock), 0, dev.stream, partials,
                       static_cast<int>(grid), total);
    PLK_HIP(hipGetLastError());
    double out = 0.0;
    PLK_HIP(hipMemcpyAsync(&out, total, sizeof(double), hipMemcpyDeviceToHost, dev.stream));
    // Also keeps b.bytes alive until the copy above has read it.
    PLK_HIP(hipStreamSynchronize(dev.stream));
    return out;
}

std::vector<unsigned char> pauli_image(const std::vector<PauliTerm> &terms, size_t &toff,
                                       int &ngroups) {
    PauliBlob b = pack_pauli(terms);
    toff = b.toff;
    ngroups = b.ngroups;
    return std::move(b.bytes);
}

void launch_pauli_apply(Device &dev, const State &out, const State &in, const PauliDevice &td) {
    launch_pauli_apply_at(dev, out, in, td.blob, td.toff, td.ngroups);
}

void launch_pauli_apply_at(Device &dev, const State &out, const State &in, const void *blob,
                           size_t toff, int ngroups) {
    if (ngroups == 0) {
        launch_fill_zero(dev, out);
        return;
    }
    const PauliTables t = view_at(blob, toff, ngroups);
    const unsigned grid = grid_for(dev, in.len);
    ProfScope prof("pauli_apply", static_cast<double>(t.ngroups + 1) * in.bytes(), dev.stream);
    if (in.dtype == PLK_C64) {
        hipLaunchKernelGGL((pauli_apply_kernel<float>), dim3(grid), dim3(kBlock), 0, dev.stream,
                           static_cast<Cx<float> *>(out.d), static_cast<const Cx<float> *>(in.d),
                           static_cast<uint64_t>(in.len), t.groups, t.ngroups, t.terms);
    } else {
        hipLaunchKernelGGL((pauli_apply_kernel<double>), dim3(grid), dim3(kBlock), 0,
                           dev.stream, static_cast<Cx<double> *>(out.d),
                           static_cast<const Cx<double> *>(in.d), static_cast<uint64_t>(in.len),
                           t.groups, t.ngroups, t.terms);
    }
    PLK_HIP(hipGetLastError());
}

void launch_pauli_apply(Device &dev, const State &out, const State &in,
                        const std::vector<PauliTerm> &terms) {
    if (terms.empty()) {
        launch_fill_zero(dev, out);
        return;
    }
    // The temporary tables are freed (after a stream sync) when this returns.
    launch_pauli_apply(dev, out, in, *pauli_upload(dev, terms));
}

CsrDevice::~CsrDevice() {
    if (blob != nullptr) {
        // Destructors do not throw: errors of the teardown are ignored.
        (void)hipSetDevice(device);
        (void)hipStreamSynchronize(stream);
        (void)hipFree(blob);
    }
}

void csr_check_shape(const CsrView &m, size_t len) {
    if (m.rows != len) {
        raise(PLK_ERR_INVALID, "CSR matrix has " + std::to_string(m.rows) +
                                   " rows but the state has " + std::to_string(len) +
                                   " amplitudes");
    }
    const int64_t first = m.idx32 ? static_cast<const int32_t *>(m.indptr)[0]
                                  : static_cast<const int64_t *>(m.indptr)[0];
    const int64_t last = m.idx32 ? static_cast<const int32_t *>(m.indptr)[m.rows]
                                 : static_cast<const int64_t *>(m.indptr)[m.rows];
    if (first != 0 || last < 0 || static_cast<size_t>(last) != m.nnz) {
        raise(PLK_ERR_INVALID, "CSR indptr does not match the number of non-zeros");
    }
}

namespace {

// Lanes per row: the power of two nearest above the average row length,
// from 4 to 64.
int csr_lanes(size_t nnz, size_t rows) {
    const char *e = std::getenv("PLK_CSR_LANES");
    if (e && *e) {
        const int g = std::atoi(e);
        if (g == 4 || g == 8 || g == 16 || g == 32 || g == 64) {
            return g;
        }
    }
    // Each lane takes four entries per round, so G lanes cover 4 G entries of
    // a row at once. Measured on a 20-qubit Hamiltonian (47 entries a row):
    // G = 4, 8, 16, 32, 64 -> 0.54, 0.35, 0.26, 0.27, 0.36 ms (complex128).
    const double avg = rows ? static_cast<double>(nnz) / static_cast<double>(rows) : 0.0;
    return avg <= 6.0 ? 4 : avg <= 12.0 ? 8 : avg <= 96.0 ? 16 : avg <= 192.0 ? 32 : 64;
}

// Grid and rows per block: contiguous ranges, a multiple of the rows one
// block iteration covers, at most 8 blocks per CU.
void csr_grid(const Device &dev, size_t rows, int G, unsigned &grid, uint64_t &chunk) {
    const uint64_t step = static_cast<uint64_t>(64 / G) * (kBlock / 64);
    const uint64_t groups = (rows + step - 1) / step;
    const uint64_t cap = static_cast<uint64_t>(dev.num_cus) * 8;
    const uint64_t blocks = std::max<uint64_t>(1, std::min(groups, cap));
    chunk = std::max<uint64_t>(1, (groups + blocks - 1) / blocks) * step;
    grid = static_cast<unsigned>(std::max<uint64_t>(1, (rows + chunk - 1) / chunk));
}

// Raw device views of one CSR launch.
struct CsrArgs {
    const void *data, *indices, *indptr;
    bool idx32, vdouble;
    size_t nnz, rows;
};

template <typename T, typename V, typename I>
void csr_launch_t(Device &dev, const CsrArgs &m, const State &in, Cx<T> *out, unsigned grid,
                  uint64_t chunk, int G, double *partials) {
    const auto *psi = static_cast<const Cx<T> *>(in.d);
    const auto *d = static_cast<const Cx<V> *>(m.data);
    const auto *ix = static_cast<const I *>(m.indices);
    const auto *ip = static_cast<const I *>(m.indptr);
    const uint64_t len = in.len, rows = m.rows, nnz = m.nnz;
#define PLK_CSR_LAUNCH(GG)                                                                        \
    if (out == nullptr) {                                                                         \
        hipLaunchKernelGGL((csr_expval_kernel<T, V, I, GG>), dim3(grid), dim3(kBlock), 0,       \
                           dev.stream, psi, len, d, ix, ip, rows, nnz, chunk, partials);          \
    } else {                                                                                      \
        hipLaunchKernelGGL((csr_spmv_kernel<T, V, I, GG>), dim3(grid), dim3(kBlock), 0,         \
                           dev.stream, out, psi, len, d, ix, ip, rows, nnz, chunk, partials);     \
    }
    if (G == 4) {
        PLK_CSR_LAUNCH(4)
    } else if (G == 8) {
        PLK_CSR_LAUNCH(8)
    } else if (G == 16) {
        PLK_CSR_LAUNCH(16)
    } else if (G == 32) {
        PLK_CSR_LAUNCH(32)
    } else {
        PLK_CSR_LAUNCH(64)
    }
#undef PLK_CSR_LAUNCH
    PLK_HIP(hipGetLastError());
}

// One CSR kernel over `in`; out == nullptr computes the expectation value.
// Returns {sum, number of blocks that met a malformed row or index}.
std::pair<double, double> csr_run(Device &dev, const CsrArgs &m, const State &in,
                                  const State *out) {
    const int G = csr_lanes(m.nnz, m.rows);
    unsigned grid = 1;
    uint64_t chunk = 1;
    csr_grid(dev, m.rows, G, grid, chunk);
    ensure_red(dev, static_cast<size_t>(grid) * 2, 2);
    {
        const double vb = m.vdouble ? 16.0 : 8.0;
        const double ib = m.idx32 ? 4.0 : 8.0;
        const double bytes = static_cast<double>(m.nnz) * (vb + ib) +
                             ib * static_cast<double>(m.rows + 1) +
                             (out ? 2.0 : 1.0) * static_cast<double>(in.bytes());
        ProfScope prof(out ? "csr_spmv" : "csr_expval", bytes, dev.stream);
        if (in.dtype == PLK_C64) {
            auto *o = out ? static_cast<Cx<float> *>(out->d) : nullptr;
            if (m.vdouble) {
                if (m.idx32) {
                    csr_launch_t<float, double, int32_t>(dev, m, in, o, grid, chunk, G, dev.red);
                } else {
                    csr_launch_t<float, double, int64_t>(dev, m, in, o, grid, chunk, G, dev.red);
                }
            } else if (m.idx32) {
                csr_launch_t<float, float, int32_t>(dev, m, in, o, grid, chunk, G, dev.red);
            } else {
                csr_launch_t<float, float, int64_t>(dev, m, in, o, grid, chunk, G, dev.red);
            }
        } else {
            if (!m.vdouble) {
                raise(PLK_ERR_INVALID, "complex128 states need FP64 CSR values");
            }
            auto *o = out ? static_cast<Cx<double> *>(out->d) : nullptr;
            if (m.idx32) {
                csr_launch_t<double, double, int32_t>(dev, m, in, o, grid, chunk, G, dev.red);
            } else {
                csr_launch_t<double, double, int64_t>(dev, m, in, o, grid, chunk, G, dev.red);
            }
        }
    }
    double res[2] = {0.0, 0.0};
    if (out == nullptr) {
        finish_sums(dev, static_cast<int>(grid), 2, res);
    } else {
        // The flags sit where the expectation kernel keeps its sums.
        finish_sums(dev, static_cast<int>(grid), 1, res + 1);
    }
    if (res[1] != 0.0) {
        raise(PLK_ERR_INVALID,
              "CSR matrix has a row range or column index outside the matrix "
              "(indptr must be non-decreasing, 0 <= indices < rows)");
    }
    return {res[0], res[1]};
}

CsrArgs args_of(const CsrDevice &m) {
    return {m.data, m.indices, m.indptr, m.idx32, m.dtype == PLK_C128, m.nnz, m.rows};
}

template <typename Src, typename Dst> void narrow_copy(const Src *src, size_t n, Dst *dst) {
    for (size_t k = 0; k < n; k++) {
        dst[k] = static_cast<Dst>(src[k]);
    }
}

} // namespace

std::unique_ptr<CsrDevice> csr_upload(Device &dev, const CsrView &m, int dtype) {
    auto c = std::make_unique<CsrDevice>();
    c->device = dev.id;
    c->stream = dev.stream;
    c->dtype = dtype;
    c->nnz = m.nnz;
    c->rows = m.rows;
    c->idx32 = m.rows <= (size_t{1} << 31) && m.nnz <= static_cast<size_t>(INT32_MAX);
    const size_t vb = dtype == PLK_C64 ? 8 : 16;
    const size_t ib = c->idx32 ? 4 : 8;
    const size_t o1 = (m.nnz * vb + 255) / 256 * 256;
    const size_t o2 = o1 + (m.nnz * ib + 255) / 256 * 256;
    const size_t total = o2 + (m.rows + 1) * ib + 16;
    // Host image in the device layout, then one copy.
    std::vector<unsigned char> img(total);
    if (dtype == PLK_C64) {
        narrow_copy(m.data, 2 * m.nnz, reinterpret_cast<float *>(img.data()));
    } else if (m.nnz) {
        std::memcpy(img.data(), m.data, m.nnz * 16);
    }
    auto put_index = [&](const void *src, size_t n, size_t off) {
        if (c->idx32 && m.idx32) {
            std::memcpy(img.data() + off, src, n * 4);
        } else if (c->idx32) {
            narrow_copy(static_cast<const int64_t *>(src), n,
                        reinterpret_cast<int32_t *>(img.data() + off));
        } else if (m.idx32) {
            narrow_copy(static_cast<const int32_t *>(src), n,
                        reinterpret_cast<int64_t *>(img.data() + off));
        } else {
            std::memcpy(img.data() + off, src, n * 8);
        }
    };
    put_index(m.indices, m.nnz, o1);
    put_index(m.indptr, m.rows + 1, o2);
    c->blob = dev_malloc(dev.id, total, "CSR matrix");
    auto *b = static_cast<unsigned char *>(c->blob);
    PLK_HIP(hipMemcpy(b, img.data(), total, hipMemcpyHostToDevice));
    c->data = b;
    c->indices = b + o1;
    c->indptr = b + o2;
    return c;
}

double csr_expval(Device &dev, const State &s, const CsrDevice &m) {
    if (m.rows != s.len) {
        raise(PLK_ERR_INVALID, "CSR matrix does not match the state's length");
    }
    return csr_run(dev, args_of(m), s, nullptr).first;
}

double csr_expval_once(Device &dev, const State &s, const CsrView &m) {
    csr_check_shape(m, s.len);
    const size_t ib = m.idx32 ? 4 : 8;
    const size_t o1 = (m.nnz * 16 + 255) / 256 * 256;
    const size_t o2 = o1 + (m.nnz * ib + 255) / 256 * 256;
    const size_t total = o2 + (m.rows + 1) * ib + 16;
    auto *b = static_cast<unsigned char *>(ensure_work(dev, total));
    if (m.nnz) {
        PLK_HIP(hipMemcpyAsync(b, m.data, m.nnz * 16, hipMemcpyHostToDevice, dev.stream));
        PLK_HIP(hipMemcpyAsync(b + o1, m.indices, m.nnz * ib, hipMemcpyHostToDevice, dev.stream));
    }
    PLK_HIP(hipMemcpyAsync(b + o2, m.indptr, (m.rows + 1) * ib, hipMemcpyHostToDevice,
                           dev.stream));
    const CsrArgs a{b, b + o1, b + o2, m.idx32, true, m.nnz, m.rows};
    return csr_run(dev, a, s, nullptr).first;
}

void csr_spmv(Device &dev, const State &out, const State &in, const CsrDevice &m) {
    if (m.rows != in.len || out.len != in.len || out.dtype != in.dtype) {
        raise(PLK_ERR_INVALID, "CSR matrix does not match the state's length");
    }
    csr_run(dev, args_of(m), in, &out);
}

} // namespace plk
