// Device context, error plumbing and kernel statistics.
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include <chrono>

#include "plk_internal.hpp"

namespace plk {

[[noreturn]] void raise(int code, const std::string &msg) { throw Error{code, msg}; }

void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory) ? PLK_ERR_OOM : PLK_ERR_HIP;
        raise(code, std::string("HIP error ") + hipGetErrorString(e) + " in " + what);
    }
}

namespace {
std::mutex g_dev_mutex;
std::map<int, std::unique_ptr<Device>> g_devices;
} // namespace

std::recursive_mutex &engine_mutex() {
    static std::recursive_mutex m;
    return m;
}

void set_device(int id) { PLK_HIP(hipSetDevice(id)); }

Device &device(int id) {
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    auto it = g_devices.find(id);
    if (it != g_devices.end()) {
        set_device(id);
        return *it->second;
    }
    int count = 0;
    PLK_HIP(hipGetDeviceCount(&count));
    if (id < 0 || id >= count) {
        raise(PLK_ERR_INVALID, "device id " + std::to_string(id) + " out of range (" +
                                   std::to_string(count) + " devices)");
    }
    auto d = std::make_unique<Device>();
    d->id = id;
    set_device(id);
    PLK_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    int cus = 0;
    PLK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, id));
    d->num_cus = cus > 0 ? cus : 256;
    Device &ref = *d;
    g_devices[id] = std::move(d);
    ensure_red(ref, static_cast<size_t>(ref.num_cus) * 8 * SEG_PMAX * SEG_NJMAX, 256);
    return ref;
}

void ensure_red(Device &d, size_t partial_doubles, size_t out_doubles) {
    if (partial_doubles > d.red_cap) {
        PLK_HIP(hipStreamSynchronize(d.stream));
        if (d.red) {
            PLK_HIP(hipFree(d.red));
            d.red = nullptr;
        }
        d.red = static_cast<double *>(dev_malloc(d.id, partial_doubles * sizeof(double), "partials"));
        d.red_cap = partial_doubles;
    }
    if (out_doubles > d.red_out_cap) {
        PLK_HIP(hipStreamSynchronize(d.stream));
        if (d.red_out) {
            PLK_HIP(hipFree(d.red_out));
            d.red_out = nullptr;
        }
        d.red_out = static_cast<double *>(dev_malloc(d.id, out_doubles * sizeof(double), "reductions"));
        d.red_out_cap = out_doubles;
    }
}

namespace {
uint64_t g_call = 1;
} // namespace

void begin_call() { g_call++; }

void *arena_alloc(Device &d, size_t bytes) {
    if (d.arena_call != g_call) {
        d.arena_call = g_call;
        d.arena_used = 0;
        if (!d.arena_retired.empty()) {
            PLK_HIP(hipStreamSynchronize(d.stream));
            for (void *p : d.arena_retired) {
                PLK_HIP(hipFree(p));
            }
            d.arena_retired.clear();
        }
    }
    const size_t need = (bytes + 255) & ~size_t{255};
    if (d.arena_used + need > d.arena_cap) {
        // Earlier regions of this call may still be read by queued kernels:
        // the outgrown chunk is retired, not freed.
        if (d.arena) {
            d.arena_retired.push_back(d.arena);
        }
        const size_t cap = std::max(need + (size_t{1} << 20), 2 * d.arena_cap);
        d.arena = dev_malloc(d.id, cap, "plan tables");
        d.arena_cap = cap;
        d.arena_used = 0;
    }
    void *p = static_cast<unsigned char *>(d.arena) + d.arena_used;
    d.arena_used += need;
    return p;
}

void *ensure_work(Device &d, size_t bytes) {
    if (bytes > d.work_cap) {
        PLK_HIP(hipStreamSynchronize(d.stream));
        if (d.work) {
            PLK_HIP(hipFree(d.work));
            d.work = nullptr;
            d.work_cap = 0;
        }
        d.work = dev_malloc(d.id, bytes, "workspace");
        d.work_cap = bytes;
    }
    return d.work;
}

double *ensure_jac(Device &d, size_t doubles) {
    if (doubles > d.jac_cap) {
        PLK_HIP(hipStreamSynchronize(d.stream));
        if (d.jac) {
            PLK_HIP(hipFree(d.jac));
            d.jac = nullptr;
            d.jac_cap = 0;
        }
        const size_t cap = std::max<size_t>(doubles, 4096);
        d.jac = static_cast<double *>(dev_malloc(d.id, cap * sizeof(double), "Jacobian"));
        d.jac_cap = cap;
    }
    return d.jac;
}

// ---------------------------------------------------------------------------
// Kernel statistics. Events are recorded around launches without blocking;
// their durations are resolved when statistics are read.
// ---------------------------------------------------------------------------
namespace {
struct Pending {
    std::string family;
    double bytes;
    hipEvent_t e0, e1;
};
std::mutex g_prof_mutex;
bool g_prof_on = false;
std::vector<Pending> g_pending;
std::map<std::string, KernelStats> g_stats;

void resolve_pending() {
    for (auto &p : g_pending) {
        PLK_HIP(hipEventSynchronize(p.e1));
        float ms = 0.f;
        PLK_HIP(hipEventElapsedTime(&ms, p.e0, p.e1));
        auto &s = g_stats[p.family];
        s.launches += 1;
        s.ms += ms;
        s.bytes += p.bytes;
        PLK_HIP(hipEventDestroy(p.e0));
        PLK_HIP(hipEventDestroy(p.e1));
    }
    g_pending.clear();
}
} // namespace

bool profiling_on() { return g_prof_on; }

namespace {
bool g_capturing = false;
} // namespace

void set_capturing(bool on) { g_capturing = on; }

ProfScope::ProfScope(const char *f, double b, hipStream_t s) : family(f), bytes(b), stream(s) {
    // Event timing does not apply to work recorded into a graph.
    if (!g_prof_on || g_capturing) {
        return;
    }
    PLK_HIP(hipEventCreate(&e0));
    PLK_HIP(hipEventCreate(&e1));
    PLK_HIP(hipEventRecord(e0, stream));
}

ProfScope::~ProfScope() {
    if (!e0) {
        return;
    }
    if (hipEventRecord(e1, stream) != hipSuccess) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return;
    }
    std::lock_guard<std::mutex> lock(g_prof_mutex);
    g_pending.push_back({family, bytes, e0, e1});
}

namespace {
double host_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
} // namespace

HostScope::HostScope(const char *f) : family(f) {
    if (g_prof_on) {
        t0 = host_ms();
    }
}

HostScope::~HostScope() {
    if (t0 >= 0.0) {
        record_kernel(family, static_cast<float>(host_ms() - t0), 0.0);
    }
}

void record_kernel(const char *family, float ms, double bytes) {
    std::lock_guard<std::mutex> lock(g_prof_mutex);
    auto &s = g_stats[family];
    s.launches += 1;
    s.ms += ms;
    s.bytes += bytes;
}

// Temporary states are recycled per device instead of being freed: every
// kernel of the library runs on the device's one stream, so a buffer handed
// out again is written only after the work that last used it. The pool holds
// at most kPoolMaxBytes per device; an allocation that fails releases it.
namespace {
constexpr size_t kPoolMaxBytes = size_t{32} << 30;

struct TmpPool {
    std::multimap<size_t, void *> free;
    size_t held = 0;
};

std::mutex g_pool_mutex;
std::map<int, TmpPool> g_pools;

void release_pool(int dev) {
    std::vector<void *> bufs;
    {
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        TmpPool &p = g_pools[dev];
        for (auto &kv : p.free) {
            bufs.push_back(kv.second);
        }
        p.free.clear();
        p.held = 0;
    }
    if (!bufs.empty()) {
        (void)hipDeviceSynchronize();
        for (void *b : bufs) {
            (void)hipFree(b);
        }
    }
}
} // namespace

void *dev_malloc(int device, size_t bytes, const char *what) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) == hipSuccess) {
        return p;
    }
    (void)hipGetLastError();
    release_pool(device);
    if (hipMalloc(&p, bytes) == hipSuccess) {
        return p;
    }
    (void)hipGetLastError();
    raise(PLK_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed (" + what + ")");
}

TmpState::TmpState(int n, int dtype, int dev) {
    s.n = n;
    s.dtype = dtype;
    s.len = size_t{1} << n;
    s.device = dev;
    const size_t bytes = s.bytes();
    {
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        TmpPool &p = g_pools[dev];
        auto it = p.free.find(bytes);
        if (it != p.free.end()) {
            s.d = it->second;
            p.held -= bytes;
            p.free.erase(it);
            return;
        }
    }
    s.d = dev_malloc(dev, bytes, "temporary state");
}

TmpState::~TmpState() {
    if (!s.d) {
        return;
    }
    const size_t bytes = s.bytes();
    {
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        TmpPool &p = g_pools[s.device];
        if (p.held + bytes <= kPoolMaxBytes) {
            p.free.emplace(bytes, s.d);
            p.held += bytes;
            return;
        }
    }
    // Kernels enqueued on this buffer must finish before it is released.
    (void)hipDeviceSynchronize();
    (void)hipFree(s.d);
}

} // namespace plk

namespace plk {

void profiling_enable_impl(bool on) {
    std::lock_guard<std::mutex> lock(g_prof_mutex);
    resolve_pending();
    g_stats.clear();
    g_prof_on = on;
}

KernelStats profiling_get_impl(const std::string &kernel) {
    std::lock_guard<std::mutex> lock(g_prof_mutex);
    resolve_pending();
    auto it = g_stats.find(kernel);
    return it == g_stats.end() ? KernelStats{} : it->second;
}

std::string profiling_names_impl() {
    std::lock_guard<std::mutex> lock(g_prof_mutex);
    resolve_pending();
    std::string s;
    for (auto &kv : g_stats) {
        if (!s.empty()) {
            s += ",";
        }
        s += kv.first;
    }
    return s;
}

} // namespace plk

namespace plk {

namespace {
template <typename T>
int launch_rtile_t(Device &dev, const State &phi, void *out, void *const *h, const void *p,
                   const RtLaunch &L) {
    switch (L.nj) {
    case 0: return rtile_launch_t<T, 0>(dev, phi, out, h, p, L);
    case 1: return rtile_launch_t<T, 1>(dev, phi, out, h, p, L);
    case 2: return rtile_launch_t<T, 2>(dev, phi, out, h, p, L);
    case 3: return rtile_launch_t<T, 3>(dev, phi, out, h, p, L);
    case 4: return rtile_launch_t<T, 4>(dev, phi, out, h, p, L);
    case 5: return rtile_launch_t<T, 5>(dev, phi, out, h, p, L);
    default: break;
    }
    raise(PLK_ERR_INVALID, "unsupported register-tile observable chunk " + std::to_string(L.nj));
}
} // namespace

int launch_rtile(Device &dev, const State &phi, const State *phi_out, void *const *h_ptrs,
                 const void *d_prog, const RtLaunch &L) {
    const int R = phi.dtype == PLK_C64 ? 4 : 3;
    const int nc = std::max(1, L.nc);
    if (L.R != R || L.T != R + RT_LANE_BITS || L.T > L.n || L.nslots > RT_PMAX || L.nj < 0 ||
        nc * std::max(1, L.nj) > RT_OBS_MAX || (nc > 1 && L.nj == 0) || L.nbig > nc ||
        (L.nj == 1 && L.nbig < nc)) {
        raise(PLK_ERR_INVALID, "register-tile pass does not match the state");
    }
    if (phi_out != nullptr && (phi_out->len != phi.len || phi_out->dtype != phi.dtype)) {
        raise(PLK_ERR_INVALID, "register-tile pass: output state of another shape");
    }
    if (nc > 1 && L.write_phi && (phi_out == nullptr || phi_out->d == phi.d)) {
        raise(PLK_ERR_INVALID, "register-tile pass: chunks side by side need a separate output");
    }
    void *out = phi_out ? phi_out->d : phi.d;
    return phi.dtype == PLK_C64 ? launch_rtile_t<float>(dev, phi, out, h_ptrs, d_prog, L)
                                : launch_rtile_t<double>(dev, phi, out, h_ptrs, d_prog, L);
}

} // namespace plk
