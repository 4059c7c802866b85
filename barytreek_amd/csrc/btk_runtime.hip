// Runtime support of libbtk: cached device allocator, scans and sorts,
// interpolation node tables, kernel parameters.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <numbers>
#include <unordered_map>
#include <vector>

#include "btk_internal.hpp"

namespace btk {

// ---------------------------------------------------------------------------
// Cached device allocator.  Blocks are kept by size and handed out again to
// requests between 1x and 2x of their size; nothing is returned to HIP until
// release_cached().  Allocations are per current device.
//
// A block freed during an API call may still be read by work queued on that
// call's stream, so until the call ends (every entry point synchronizes its
// stream before it returns) the block is handed out only to the same thread,
// whose later work is queued behind it on the same stream.  end_call() makes
// the thread's freed blocks available to every thread.
// ---------------------------------------------------------------------------
namespace {
struct Block {
    size_t bytes;
    int device;
    int owner;  // thread that freed it during a call that has not ended; 0 = any
};
std::mutex g_mu;
std::multimap<std::pair<int, size_t>, void *> g_free;  // (device, bytes) -> ptr
std::unordered_map<void *, Block> g_used;
std::unordered_map<void *, Block> g_cached;
std::atomic<int> g_next_thread{1};
thread_local int t_thread = 0;
thread_local std::vector<void *> t_freed;

int this_thread() {
    if (!t_thread) t_thread = g_next_thread.fetch_add(1);
    return t_thread;
}

size_t round_bytes(size_t b) {
    if (b < 256) return 256;
    return (b + 255) & ~size_t(255);
}
}  // namespace

void *dmalloc(size_t bytes) {
    int dev = 0;
    BTK_HIP(hipGetDevice(&dev));
    size_t b = round_bytes(bytes);
    const int me = this_thread();
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (auto it = g_free.lower_bound({dev, b});
             it != g_free.end() && it->first.first == dev && it->first.second <= 2 * b; ++it) {
            void *p = it->second;
            const Block blk = g_cached[p];
            if (blk.owner != 0 && blk.owner != me) continue;
            g_free.erase(it);
            g_cached.erase(p);
            g_used[p] = Block{blk.bytes, dev, 0};
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        // drop the cache once and retry
        release_cached();
        BTK_HIP(hipMalloc(&p, b));
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_used[p] = Block{b, dev, 0};
    return p;
}

void dfree(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_used.find(p);
    if (it == g_used.end()) return;
    Block blk = it->second;
    g_used.erase(it);
    blk.owner = this_thread();
    g_free.emplace(std::make_pair(blk.device, blk.bytes), p);
    g_cached[p] = blk;
    t_freed.push_back(p);
}

void end_call() {
    std::lock_guard<std::mutex> lk(g_mu);
    for (void *p : t_freed) {
        auto it = g_cached.find(p);
        if (it != g_cached.end() && it->second.owner == t_thread) it->second.owner = 0;
    }
    t_freed.clear();
}

void release_cached() {
    std::lock_guard<std::mutex> lk(g_mu);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto &kv : g_free) {
        if (kv.first.first != cur) (void)hipSetDevice(kv.first.first);
        (void)hipFree(kv.second);
        if (kv.first.first != cur) (void)hipSetDevice(cur);
    }
    g_free.clear();
    g_cached.clear();
}

namespace {
__global__ void k_zero(size_t n, double *__restrict__ p) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}
}  // namespace

void zero_fill(double *p, size_t n, hipStream_t s) {
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    k_zero<<<(unsigned)blocks, 256, 0, s>>>(n, p);
    BTK_HIP(hipGetLastError());
}

HostScratch::HostScratch() { BTK_HIP(hipHostMalloc(reinterpret_cast<void **>(&v), 64 * sizeof(long long))); }
HostScratch::~HostScratch() {
    if (v) (void)hipHostFree(v);
}
HostScratch &host_scratch() {
    thread_local HostScratch hs;
    return hs;
}

// ---------------------------------------------------------------------------
// Scans and sorts.
// ---------------------------------------------------------------------------
void inclusive_sum_i64(const long long *in, long long *out, int n, hipStream_t s) {
    if (n <= 0) return;
    size_t tmp = 0;
    BTK_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, in, out, n, s));
    DBuf<char> t(tmp ? tmp : 1);
    BTK_HIP(hipcub::DeviceScan::InclusiveSum(t.p, tmp, in, out, n, s));
}

void inclusive_sum_i32(const int *in, int *out, int n, hipStream_t s) {
    if (n <= 0) return;
    size_t tmp = 0;
    BTK_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, in, out, n, s));
    DBuf<char> t(tmp ? tmp : 1);
    BTK_HIP(hipcub::DeviceScan::InclusiveSum(t.p, tmp, in, out, n, s));
}

void exclusive_sum_i32(const int *in, int *out, int n, hipStream_t s) {
    if (n <= 0) return;
    size_t tmp = 0;
    BTK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, s));
    DBuf<char> t(tmp ? tmp : 1);
    BTK_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, in, out, n, s));
}

void sort_pairs_i32(const int *keys_in, int *keys_out, const int *vals_in, int *vals_out, int n, int key_bits,
                    hipStream_t s) {
    if (n <= 0) return;
    size_t tmp = 0;
    BTK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys_in, keys_out, vals_in, vals_out, n, 0, key_bits,
                                               s));
    DBuf<char> t(tmp ? tmp : 1);
    BTK_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, keys_in, keys_out, vals_in, vals_out, n, 0, key_bits, s));
}

void sort_pairs_u64(const unsigned long long *keys_in, unsigned long long *keys_out, const int *vals_in,
                    int *vals_out, int n, int key_bits, hipStream_t s) {
    if (n <= 0) return;
    size_t tmp = 0;
    BTK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys_in, keys_out, vals_in, vals_out, n, 0, key_bits,
                                               s));
    DBuf<char> t(tmp ? tmp : 1);
    BTK_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, keys_in, keys_out, vals_in, vals_out, n, 0, key_bits, s));
}

// ---------------------------------------------------------------------------
// Node tables: the same expressions as the reference, evaluated by host libm.
// ---------------------------------------------------------------------------
NodeTables node_tables(int n) {
    NodeTables t{};
    const double pi = std::numbers::pi_v<double>;
    for (int i = 0; i <= n && i <= BTK_MAX_DEGREE; i++) {
        t.a[i] = std::cos(pi * i / n);  // bli_impl.hpp:25
        t.b[i] = std::cos(pi / n * i);  // bli_impl.hpp:74
    }
    return t;
}

int components(int dim, int kernel) {
    if (kernel == BTK_KERNEL_POT) return 1;
    if (kernel == BTK_KERNEL_REF_VEL3D) return 3;
    return 1 + dim;
}

KParams make_kparams(int dim, double eps, const btk_options &o) {
    KParams k{};
    k.dim = dim;
    k.kernel = o.kernel;
    k.excl = o.exclude_self;
    k.eps2 = eps * eps;
    const double pi = std::numbers::pi_v<double>;
    // poisson_fmm_impl.hpp:38 (2D) and :241 (3D)
    double g = dim == 2 ? 1.0 / (4.0 * pi) : -1.0 / (4.0 * pi);
    k.gfc = o.convention ? -g : g;
    return k;
}

}  // namespace btk
