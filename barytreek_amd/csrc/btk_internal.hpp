// Internal declarations shared by the HIP translation units of libbtk.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/btk.h"

namespace btk {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define BTK_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::btk::Error(std::string(#call) + " failed: " + hipGetErrorString(e_) + " (" +   \
                               __FILE__ + ":" + std::to_string(__LINE__) + ")");                   \
    } while (0)

#define BTK_REQUIRE(cond, msg)                                 \
    do {                                                       \
        if (!(cond)) throw ::btk::Error(std::string(msg));     \
    } while (0)

#define BTK_LAUNCH_CHECK() BTK_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Device memory: a size-bucketed cache in front of hipMalloc, so that
// repeated solves of the same shape allocate nothing after the first.
// No DBuf may live in static storage: it would be freed at exit, after the
// cache's thread-local bookkeeping has been destroyed.
// ---------------------------------------------------------------------------
void *dmalloc(size_t bytes);
void dfree(void *p);
void release_cached();
// Called when an API call ends, after its stream has been synchronized.
void end_call();

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { resize(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            reset();
            p = o.p; n = o.n;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() {
        if (p) dfree(p);
        p = nullptr;
        n = 0;
    }
    // discards contents
    void resize(size_t count) {
        if (count == n && p) return;
        reset();
        if (count) p = static_cast<T *>(dmalloc(count * sizeof(T)));
        n = count;
    }
    // keeps contents (first min(n, count) elements).  The old block goes back
    // to the cache with the copy still queued on s; the cache gives it only to
    // this thread's later work (queued behind the copy) until the call ends.
    void grow(size_t count, hipStream_t s) {
        if (count <= n) return;
        T *q = static_cast<T *>(dmalloc(count * sizeof(T)));
        if (p && n) BTK_HIP(hipMemcpyAsync(q, p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        reset();
        p = q;
        n = count;
    }
    T *get() const { return p; }
};

// Zeroes n doubles with a kernel (unlike hipMemsetAsync, safe to capture in
// a HIP graph that is replayed).
void zero_fill(double *p, size_t n, hipStream_t s);

// Small pinned host staging buffer for scalar read-backs.
struct HostScratch {
    long long *v = nullptr;
    HostScratch();
    ~HostScratch();
};
HostScratch &host_scratch();

// ---------------------------------------------------------------------------
// Device-side scans and sorts (rocPRIM/hipCUB), temp storage from the cache.
// ---------------------------------------------------------------------------
void inclusive_sum_i64(const long long *in, long long *out, int n, hipStream_t s);
void inclusive_sum_i32(const int *in, int *out, int n, hipStream_t s);
void exclusive_sum_i32(const int *in, int *out, int n, hipStream_t s);
// Stable sort of (key, value) pairs by key; keys are < (1 << key_bits).
void sort_pairs_u64(const unsigned long long *keys_in, unsigned long long *keys_out, const int *vals_in,
                    int *vals_out, int n, int key_bits, hipStream_t s);
void sort_pairs_i32(const int *keys_in, int *keys_out, const int *vals_in, int *vals_out, int n, int key_bits,
                    hipStream_t s);

// ---------------------------------------------------------------------------
// Tree: panels as SoA in HBM, points sorted leaf by leaf.
// ---------------------------------------------------------------------------
struct Tree {
    int dim = 3;
    int n = 0;
    int thresh = 0;
    int nprob = 1;   // independent problems: panels 0..nprob-1 are their roots (level 0)
    int pc = 0;      // panel count
    int levels = 0;
    int cap = 0;     // panel capacity of the device arrays
    std::vector<int> level_start;  // host copy, [levels + 1]
    DBuf<double> lo[3], hi[3];
    DBuf<double> radius;
    DBuf<int> count, start, parent, child0, level;
    DBuf<int> perm;        // sorted position -> original point index
    DBuf<int> point_leaf;  // original point index -> leaf panel id
    DBuf<double> sx, sy, sz;  // coordinates in sorted order
    // leaves, by level-major id order, for per-leaf kernels
    DBuf<int> leaves;
    int n_leaves = 0;
    int n_leaves_deepest = 0;  // the last n_leaves_deepest entries of `leaves` are on the deepest level
    int max_leaf_count = 0;
};

// ---------------------------------------------------------------------------
// Interaction list in FIFO order, plus its target-grouped forms.
// ---------------------------------------------------------------------------
struct List {
    int dim = 3;
    int count = 0;
    long long counts[4] = {0, 0, 0, 0};
    DBuf<btk_interact_pair> pairs;  // FIFO order
    // Interpolation degrees in use, ascending: deg[ndeg - 1] is the
    // configured one.  An entry's interact_type holds its type in bits 0-1
    // and its degree index in bits 2-3; one degree unless the learned MAC
    // chose per pair.
    int ndeg = 1;
    int deg[4] = {0, 0, 0, 0};
    // Entries grouped by (near/far, degree index, target panel), stable in
    // FIFO order; group g = far * ndeg + di.  "Near" entries (PP, PC: the
    // targets are points of a target leaf) and "far" entries (CP, CC: the
    // targets are the proxy points of a target cluster) of group g and panel
    // p are [goff[g*pct+p], goff[g*pct+p+1]).  PP entries have degree index 0.
    int pct = 0;
    DBuf<int> goff;             // [2*ndeg*pct + 1]
    DBuf<int> gsrc;             // source panel of each entry
    DBuf<unsigned char> gtype;  // interact_type of each entry
    double evals[4] = {0, 0, 0, 0};  // kernel evaluations per type
};

// Element offset of degree index di's block in a buffer that holds the proxy
// arrays of every degree the list uses, `comps` arrays of pc * (deg+1)^dim
// per degree; di = l.ndeg gives the buffer's size.
size_t proxy_offset(const List &l, int pc, int comps, int di);

// Interpolation tables (host-computed with libm, so that the nodes are the
// same bits as the reference's cos() calls).
struct NodeTables {
    double a[BTK_MAX_DEGREE + 1];  // cos(pi * i / n)  (bli_impl.hpp:25)
    double b[BTK_MAX_DEGREE + 1];  // cos(pi / n * i)  (bli_impl.hpp:74)
};
NodeTables node_tables(int n);

// Kernel-side parameters.
struct KParams {
    int dim;
    int kernel;
    int excl;
    double eps2;
    double gfc;
};
KParams make_kparams(int dim, double eps, const btk_options &o);
int components(int dim, int kernel);

// Process-wide launch shapes of the interaction kernels (zero = default).
btk_tuning &tuning();
// charge vectors per workgroup of the many-vector MFMA kernels (tuning, default 64)
inline int multi_vectors() { return tuning().multi_vectors ? tuning().multi_vectors : 64; }

// Phases.
std::unique_ptr<Tree> build_tree(int dim, int thresh, const double *x, const double *y, const double *z, int n,
                                 hipStream_t s);
// A forest: problem p owns points [offsets[p], offsets[p+1]) and root panel p;
// each problem's tree is the tree build_tree gives for its points alone, with
// panel ids interleaved level by level across the problems.
std::unique_ptr<Tree> build_forest(int dim, int thresh, const double *x, const double *y, const double *z,
                                   const std::vector<int> &offsets, hipStream_t s);
// learned_mac: consult the installed learned MAC where the theta MAC would
// refine (3D).  visited: when given, receives every pair the traversal visits.
std::unique_ptr<List> build_list(btk_run_config *cfg, const Tree &t, const Tree &s, hipStream_t st,
                                 bool learned_mac = false, std::vector<int2> *visited = nullptr);
void upward_pass(const btk_run_config &cfg, const Tree &src, const double *q_sorted, double *W, hipStream_t s);
void gather_sorted(const Tree &t, const double *orig, double *sorted, hipStream_t s);
// Optional HIP events recorded around the near (PP+PC) and far (CP+CC)
// kernel launches: ev[0], ev[1] bracket the near kernel, ev[2], ev[3] the far.
struct KernelEvents {
    hipEvent_t ev[4];
    bool near_done = false, far_done = false;
    long long cc_shared = 0;  // CC entries the launches evaluated as shared-matrix products
};
// W and Psi hold every degree of the list (proxy_offset with 1 and C
// components); with one degree they are the plain arrays.
// sorted_out: the near-field results go to out in the target tree's sorted
// order (coalesced adds; add_unsorted moves them to the original order).
void fmm_interactions(const btk_run_config &cfg, const btk_options &o, const Tree &tt, const Tree &ts,
                      const List &l, const double *q_sorted, const double *W, double *out, double *Psi,
                      hipStream_t s, KernelEvents *kev = nullptr, bool sorted_out = false);
// The same in two steps: the plan (descriptors and work items; allocates and
// reads counts back) and the launches alone.  The plan refers to tt, ts.
struct InteractPlan;
std::shared_ptr<InteractPlan> plan_interactions(const btk_run_config &cfg, const btk_options &o, const Tree &tt,
                                                const Tree &ts, const List &l, hipStream_t s,
                                                bool sorted_out = false);
void launch_interactions(const InteractPlan &p, const double *q_sorted, const double *W, double *out, double *Psi,
                         hipStream_t s, KernelEvents *kev = nullptr);
void downward_pass(const btk_run_config &cfg, const btk_options &o, const Tree &tt, double *Psi, double *out,
                   hipStream_t s, bool sorted_out = false);
// out[c][perm[i]] += sorted[c][i] over the tree's n points and C components.
void add_unsorted(const Tree &t, int C, const double *sorted, double *out, hipStream_t s);
// The passes over every degree of a list.  upward_all runs the upward pass at
// the configured degree and restricts its charges to each lower degree, which
// equals that degree's own upward pass: degree-d basis functions are
// polynomials the degree-n interpolation reproduces.  downward_all extends
// each lower degree's potentials into the configured degree's (the degree-d
// interpolant evaluated at the degree-n proxies, reproduced by degree-n
// interpolation in turn) and runs one downward pass.
void upward_all(const btk_run_config &cfg, const List &l, const Tree &src, const double *q_sorted, double *W,
                hipStream_t s);
void downward_all(const btk_run_config &cfg, const btk_options &o, const List &l, const Tree &tt, double *Psi,
                  double *out, hipStream_t s, bool sorted_out = false);

// Many charge vectors (btk_mrhs.hip).  q is [K][n] in original order;
// qs is [n][kp] FP32 in sorted order, kp = K rounded up to 64, zero padded.
void gather_multi_f32(const Tree &t, const double *q, int K, int kp, float *qs, hipStream_t s);
// W [K][rows] FP64 -> Wt [rows][kp] FP32
void transpose_weights_f32(long long rows, int K, int kp, const double *W, float *Wt, hipStream_t s);
// FP32 MFMA interactions of K vectors: adds PP+PC into out [K][C][M] and
// CP+CC into Psi [K][C][pc_t][P].
void fmm_interactions_multi(const btk_run_config &cfg, const btk_options &o, const Tree &tt, const Tree &ts,
                            const List &l, const float *qs, const float *wt, int K, int kp, double *out, double *Psi,
                            hipStream_t s, KernelEvents *kev);
void direct_sum(int dim, const btk_options &o, double eps, int m, const int *idx, const double *xt,
                const double *yt, const double *zt, int n, const double *xs, const double *ys,
                const double *zs, const double *q, double *out, hipStream_t s);

// Learned MAC (btk_mac.hip): the process-wide model and its training.
struct MacModel {
    DBuf<float> w;  // BTK_MAC_WEIGHTS parameters
    bool ready = false;
    double budget = 1e-7;
    float cost[4] = {1.0f, 1.0f, 1.0f, 1.0f};  // relative cost of one PP, PC, CP, CC evaluation
};
MacModel &mac_model();
void mac_set_weights(const float *h, hipStream_t s);
void mac_get_weights(float *h, hipStream_t s);
int mac_sample(const btk_run_config &cfg, int n, const double *x, const double *y, const double *z, int n_samples,
               unsigned seed, float *X, float *Y, int *pairs, hipStream_t s);
float mac_train(const float *X, const float *Y, int n, int epochs, int batch, float lr, unsigned seed, bool reinit,
                hipStream_t s);
void mac_forward(const float *X, int n, float *out, hipStream_t s);
float mac_gradients(const float *X, const float *Y, int n, float *grad, hipStream_t s);

inline int ipow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; i++) r *= b;
    return r;
}

}  // namespace btk
