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
    double acc[1] = {0.0};
    for (uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; i < len;
         i += uint64_t{gridDim.x} * blockDim.x) {
        const Cx<T> a = ld(psi, i);
        const double ar = a.r, ai = a.i;
        for (int gi = 0; gi < ngroups; gi++) {
            const PauliGroup g = groups[gi];
            const Cx<T> b = g.x ? ld(psi, i ^ g.x) : a;
            double cr, ci;
            group_coeff(terms, g, i, cr, ci);
            // Re(conj(a) * c * b)
            const double br = b.r, bi = b.i;
            const double tr = cr * br - ci * bi;
            const double ti = cr * bi + ci * br;
            acc[0] = fma(ar, tr, fma(ai, ti, acc[0]));
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

template <typename T>
__global__ void __launch_bounds__(kBlock)
    csr_expval_kernel(const Cx<T> *__restrict__ psi, const double *__restrict__ data,
                      const int64_t *__restrict__ indices, const int64_t *__restrict__ indptr,
                      uint64_t rows, double *__restrict__ partials) {
    double acc[1] = {0.0};
    for (uint64_t r = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; r < rows;
         r += uint64_t{gridDim.x} * blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int64_t k = indptr[r]; k < indptr[r + 1]; k++) {
            const Cx<T> b = ld(psi, static_cast<uint64_t>(indices[k]));
            const double dr = data[2 * k], di = data[2 * k + 1];
            sr += dr * b.r - di * b.i;
            si += dr * b.i + di * b.r;
        }
        const Cx<T> a = ld(psi, r);
        acc[0] += static_cast<double>(a.r) * sr + static_cast<double>(a.i) * si;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = acc[0];
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_kernel(Cx<T> *__restrict__ out, const Cx<T> *__restrict__ psi,
                    const double *__restrict__ data, const int64_t *__restrict__ indices,
                    const int64_t *__restrict__ indptr, uint64_t rows) {
    for (uint64_t r = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x; r < rows;
         r += uint64_t{gridDim.x} * blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int64_t k = indptr[r]; k < indptr[r + 1]; k++) {
            const Cx<T> b = ld(psi, static_cast<uint64_t>(indices[k]));
            const double dr = data[2 * k], di = data[2 * k + 1];
            sr += dr * b.r - di * b.i;
            si += dr * b.i + di * b.r;
        }
        st(out, r, Cx<T>{static_cast<T>(sr), static_cast<T>(si)});
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

PauliTables upload_pauli(Device &dev, const std::vector<PauliTerm> &terms) {
    const PauliBlob b = pack_pauli(terms);
    void *blob = nullptr;
    PLK_HIP(hipMalloc(&blob, b.bytes.size()));
    PLK_HIP(hipMemcpy(blob, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice));
    return tables_at(blob, b);
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

void free_tables(Device &dev, void *blob) {
    PLK_HIP(hipStreamSynchronize(dev.stream));
    PLK_HIP(hipFree(blob));
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

double pauli_expval(Device &dev, const State &s, const std::vector<PauliTerm> &terms) {
    if (terms.empty()) {
        return 0.0;
    }
    PauliTables t = upload_pauli(dev, terms);
    const unsigned grid = grid_for(dev, s.len);
    ensure_red(dev, grid, 1);
    launch_pauli_expval(dev, s, t, grid, dev.red);
    double out = 0.0;
    finish_sums(dev, static_cast<int>(grid), 1, &out);
    free_tables(dev, t.blob);
    return out;
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

void launch_pauli_apply(Device &dev, const State &out, const State &in,
                        const std::vector<PauliTerm> &terms) {
    if (terms.empty()) {
        launch_fill_zero(dev, out);
        return;
    }
    PauliTables t = upload_pauli(dev, terms);
    const unsigned grid = grid_for(dev, in.len);
    {
        ProfScope prof("pauli_apply", static_cast<double>(t.ngroups + 1) * in.bytes(),
                       dev.stream);
        if (in.dtype == PLK_C64) {
            hipLaunchKernelGGL((pauli_apply_kernel<float>), dim3(grid), dim3(kBlock), 0,
                               dev.stream, static_cast<Cx<float> *>(out.d),
                               static_cast<const Cx<float> *>(in.d),
                               static_cast<uint64_t>(in.len), t.groups, t.ngroups, t.terms);
        } else {
            hipLaunchKernelGGL((pauli_apply_kernel<double>), dim3(grid), dim3(kBlock), 0,
                               dev.stream, static_cast<Cx<double> *>(out.d),
                               static_cast<const Cx<double> *>(in.d),
                               static_cast<uint64_t>(in.len), t.groups, t.ngroups, t.terms);
        }
        PLK_HIP(hipGetLastError());
    }
    free_tables(dev, t.blob);
}

namespace {

struct CsrDev {
    void *blob = nullptr;
    const double *data = nullptr;
    const int64_t *indices = nullptr;
    const int64_t *indptr = nullptr;
};

CsrDev upload_csr(const State &s, const double *data, const int64_t *indices,
                  const int64_t *indptr, size_t nnz, size_t rows) {
    if (rows != s.len) {
        raise(PLK_ERR_INVALID, "CSR matrix has " + std::to_string(rows) +
                                   " rows but the state has " + std::to_string(s.len) +
                                   " amplitudes");
    }
    if (static_cast<size_t>(indptr[rows]) != nnz || indptr[0] != 0) {
        raise(PLK_ERR_INVALID, "CSR indptr does not match the number of non-zeros");
    }
    for (size_t k = 0; k < nnz; k++) {
        if (indices[k] < 0 || static_cast<size_t>(indices[k]) >= s.len) {
            raise(PLK_ERR_INVALID, "CSR column index out of range");
        }
    }
    for (size_t r = 0; r < rows; r++) {
        if (indptr[r + 1] < indptr[r]) {
            raise(PLK_ERR_INVALID, "CSR indptr must be non-decreasing");
        }
    }
    const size_t db = nnz * 2 * sizeof(double);
    const size_t ib = nnz * sizeof(int64_t);
    const size_t pb = (rows + 1) * sizeof(int64_t);
    const size_t o1 = (db + 255) / 256 * 256;
    const size_t o2 = o1 + (ib + 255) / 256 * 256;
    CsrDev c;
    PLK_HIP(hipMalloc(&c.blob, o2 + pb + 16));
    auto *b = static_cast<unsigned char *>(c.blob);
    if (nnz) {
        PLK_HIP(hipMemcpy(b, data, db, hipMemcpyHostToDevice));
        PLK_HIP(hipMemcpy(b + o1, indices, ib, hipMemcpyHostToDevice));
    }
    PLK_HIP(hipMemcpy(b + o2, indptr, pb, hipMemcpyHostToDevice));
    c.data = reinterpret_cast<const double *>(b);
    c.indices = reinterpret_cast<const int64_t *>(b + o1);
    c.indptr = reinterpret_cast<const int64_t *>(b + o2);
    return c;
}

} // namespace

double csr_expval(Device &dev, const State &s, const double *data, const int64_t *indices,
                  const int64_t *indptr, size_t nnz, size_t rows) {
    CsrDev c = upload_csr(s, data, indices, indptr, nnz, rows);
    const unsigned grid = grid_for(dev, rows);
    ensure_red(dev, grid, 1);
    {
        ProfScope prof("csr_expval",
                       static_cast<double>(nnz) * 24.0 + 8.0 * (rows + 1) + 2.0 * s.bytes(),
                       dev.stream);
        if (s.dtype == PLK_C64) {
            hipLaunchKernelGGL((csr_expval_kernel<float>), dim3(grid), dim3(kBlock), 0,
                               dev.stream, static_cast<const Cx<float> *>(s.d), c.data,
                               c.indices, c.indptr, static_cast<uint64_t>(rows), dev.red);
        } else {
            hipLaunchKernelGGL((csr_expval_kernel<double>), dim3(grid), dim3(kBlock), 0,
                               dev.stream, static_cast<const Cx<double> *>(s.d), c.data,
                               c.indices, c.indptr, static_cast<uint64_t>(rows), dev.red);
        }
        PLK_HIP(hipGetLastError());
    }
    double out = 0.0;
    finish_sums(dev, static_cast<int>(grid), 1, &out);
    free_tables(dev, c.blob);
    return out;
}

void csr_spmv(Device &dev, const State &out, const State &in, const double *data,
              const int64_t *indices, const int64_t *indptr, size_t nnz, size_t rows) {
    CsrDev c = upload_csr(in, data, indices, indptr, nnz, rows);
    const unsigned grid = grid_for(dev, rows);
    {
        ProfScope prof("csr_spmv",
                       static_cast<double>(nnz) * 24.0 + 8.0 * (rows + 1) + 2.0 * in.bytes(),
                       dev.stream);
        if (in.dtype == PLK_C64) {
            hipLaunchKernelGGL((csr_spmv_kernel<float>), dim3(grid), dim3(kBlock), 0, dev.stream,
                               static_cast<Cx<float> *>(out.d),
                               static_cast<const Cx<float> *>(in.d), c.data, c.indices,
                               c.indptr, static_cast<uint64_t>(rows));
        } else {
            hipLaunchKernelGGL((csr_spmv_kernel<double>), dim3(grid), dim3(kBlock), 0,
                               dev.stream, static_cast<Cx<double> *>(out.d),
                               static_cast<const Cx<double> *>(in.d), c.data, c.indices,
                               c.indptr, static_cast<uint64_t>(rows));
        }
        PLK_HIP(hipGetLastError());
    }
    free_tables(dev, c.blob);
}

} // namespace plk
