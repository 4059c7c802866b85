// Tree construction on the GPU (the reference's blfmm_tree_construction_2d/3d,
// tree_construction.cpp:5-306, is a host-serial BFS).
//
// The BFS is run level by level.  Every panel owns a contiguous range of the
// sorted point arrays; refining a panel is a stable partition of its range
// into 2^dim child ranges, in child order c1..c8, so each leaf ends up holding
// its points in ascending original order -- the order of the reference's
// temp_points_inside lists.  Panel ids, boxes, radii and level_start come out
// bit for bit as the reference's: children are numbered in the order the
// reference appends them (parents in id order, c1..c8), boxes are exact
// halves, and the radius is evaluated with the reference's expression (this
// file is compiled with -ffp-contract=off).
#include <cfloat>
#include <climits>

#include "btk_device.hpp"

namespace btk {
namespace {

constexpr int TB = 256;
constexpr int CHUNK = 4096;  // points per work chunk of the split kernels
constexpr int kMaxLevels = 200;

// ---------------------------------------------------------------------------
// Root box (tree_construction.cpp:145-167): min starts at DBL_MAX and max at
// DBL_MIN (numeric_limits<double>::min(), a positive number); std::min keeps
// the first of equal values, which the (value, index) reduction reproduces.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool better_min(double v, int i, double bv, int bi) {
    return v < bv || (v == bv && i < bi);
}
__device__ __forceinline__ bool better_max(double v, int i, double bv, int bi) {
    return v > bv || (v == bv && i < bi);
}

struct BoxAcc {
    double mn[3], mx[3];
    int imn[3], imx[3];
};

__device__ void box_init(BoxAcc &a) {
    for (int d = 0; d < 3; d++) {
        a.mn[d] = DBL_MAX;
        a.imn[d] = INT_MAX;
        a.mx[d] = -DBL_MAX;
        a.imx[d] = INT_MAX;
    }
}

__device__ void box_merge(BoxAcc &a, const BoxAcc &b) {
    for (int d = 0; d < 3; d++) {
        if (better_min(b.mn[d], b.imn[d], a.mn[d], a.imn[d])) {
            a.mn[d] = b.mn[d];
            a.imn[d] = b.imn[d];
        }
        if (better_max(b.mx[d], b.imx[d], a.mx[d], a.imx[d])) {
            a.mx[d] = b.mx[d];
            a.imx[d] = b.imx[d];
        }
    }
}

__device__ void block_reduce_box(BoxAcc &a, BoxAcc *sh) {
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            BoxAcc b = sh[threadIdx.x + s];
            BoxAcc c = sh[threadIdx.x];
            box_merge(c, b);
            sh[threadIdx.x] = c;
        }
        __syncthreads();
    }
    a = sh[0];
}

// Chunk c covers points [cb[c], ce[c]) of one problem.
__global__ void __launch_bounds__(TB) k_box_partial(int dim, const double *__restrict__ x, const double *__restrict__ y,
                                                    const double *__restrict__ z, const int *__restrict__ cb,
                                                    const int *__restrict__ ce, BoxAcc *__restrict__ part) {
    __shared__ BoxAcc sh[TB];
    BoxAcc a;
    box_init(a);
    const double *co[3] = {x, y, z};
    for (int i = cb[blockIdx.x] + threadIdx.x; i < ce[blockIdx.x]; i += TB) {
        for (int d = 0; d < dim; d++) {
            double v = co[d][i];
            if (better_min(v, i, a.mn[d], a.imn[d])) {
                a.mn[d] = v;
                a.imn[d] = i;
            }
            if (better_max(v, i, a.mx[d], a.imx[d])) {
                a.mx[d] = v;
                a.imx[d] = i;
            }
        }
    }
    block_reduce_box(a, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// Root panel p of the forest (one per problem) from chunks [pc0[p], pc0[p+1]).
__global__ void __launch_bounds__(TB) k_box_final(int dim, const BoxAcc *__restrict__ part,
                                                  const int *__restrict__ pc0, const int *__restrict__ off,
                                                  double *lo0, double *lo1, double *lo2, double *hi0, double *hi1,
                                                  double *hi2, int *count, int *start, int *parent, int *child0,
                                                  int *level) {
    __shared__ BoxAcc sh[TB];
    const int p = blockIdx.x;
    BoxAcc a;
    box_init(a);
    for (int i = pc0[p] + threadIdx.x; i < pc0[p + 1]; i += TB) box_merge(a, part[i]);
    block_reduce_box(a, sh);
    if (threadIdx.x == 0) {
        double *lo[3] = {lo0, lo1, lo2};
        double *hi[3] = {hi0, hi1, hi2};
        for (int d = 0; d < 3; d++) {
            if (d < dim) {
                lo[d][p] = a.mn[d];                                   // DBL_MAX when no point
                hi[d][p] = (DBL_MIN < a.mx[d]) ? a.mx[d] : DBL_MIN;  // std::max(DBL_MIN, ...)
            } else {
                lo[d][p] = 0.0;
                hi[d][p] = 0.0;
            }
        }
        count[p] = off[p + 1] - off[p];
        start[p] = off[p];
        parent[p] = -1;
        child0[p] = -1;
        level[p] = 0;
    }
}

// ---------------------------------------------------------------------------
// Level step.
// ---------------------------------------------------------------------------
// flag | (chunks << 32) for each panel of the level
__global__ void k_level_flags(int ls, int np, int thresh, const int *__restrict__ count, long long *__restrict__ val) {
    int i = blockIdx.x * TB + threadIdx.x;
    if (i >= np) return;
    int c = count[ls + i];
    long long f = c > thresh ? 1 : 0;  // tree_construction.cpp:172 (all panels of the level are leaves here)
    long long nch = f ? (c + CHUNK - 1) / CHUNK : 0;
    val[i] = f | (nch << 32);
}

// The same flags, their inclusive scan and its total (written to host memory)
// in one workgroup, for levels of at most kLevelFuse panels: one launch and
// one readback instead of the flags, a device-wide scan and a copy.
constexpr int kLevelFuse = 8192;
constexpr int kLevelThreads = 1024;
__global__ void __launch_bounds__(kLevelThreads)
    k_level_scan(int ls, int np, int thresh, const int *__restrict__ count, long long *__restrict__ val,
                 long long *__restrict__ inc, long long *host_total) {
    __shared__ long long wsum[kLevelThreads / 64];
    __shared__ long long carry;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < np; base += kLevelThreads) {
        const int i = base + tid;
        long long v = 0;
        if (i < np) {
            const int c = count[ls + i];
            const long long f = c > thresh ? 1 : 0;
            v = f | ((f ? (long long)((c + CHUNK - 1) / CHUNK) : 0) << 32);
            val[i] = v;
        }
        long long x = v;  // inclusive scan over the wavefront
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const long long y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        long long pre = carry;
        for (int k = 0; k < w; k++) pre += wsum[k];
        if (i < np) inc[i] = pre + x;
        __syncthreads();  // every thread has read carry and wsum
        if (tid == kLevelThreads - 1) carry = pre + x;
        __syncthreads();
    }
    if (tid == 0) *host_total = carry;
}

// Child boxes (tree_construction.cpp:184-194): "upper" halves are [mid, max].
__constant__ int c_hx[8] = {1, 0, 0, 1, 1, 0, 0, 1};
__constant__ int c_hy[8] = {1, 1, 0, 0, 1, 1, 0, 0};
__constant__ int c_hz[8] = {1, 1, 1, 1, 0, 0, 0, 0};

__global__ void k_make_children(int dim, int ls, int np, int pc, int lvl, const long long *__restrict__ val,
                                const long long *__restrict__ inc, double *lo0, double *lo1, double *lo2,
                                double *hi0, double *hi1, double *hi2, int *count, int *start, int *parent,
                                int *child0, int *level, int *__restrict__ chunk_panel, int *__restrict__ chunk_j) {
    int i = blockIdx.x * TB + threadIdx.x;
    if (i >= np) return;
    long long v = val[i];
    if (!(v & 1)) return;
    long long ex = inc[i] - v;
    int rank = (int)(ex & 0xffffffffLL);
    int cb = (int)(ex >> 32);
    int nch = (int)(v >> 32);
    const int nc = dim == 3 ? 8 : 4;
    const int p = ls + i;
    const int base = pc + nc * rank;
    child0[p] = base;
    double *lo[3] = {lo0, lo1, lo2};
    double *hi[3] = {hi0, hi1, hi2};
    double m1[3], m2[3], m3[3];
    for (int d = 0; d < 3; d++) {
        m1[d] = lo[d][p];
        m3[d] = hi[d][p];
        m2[d] = 0.5 * (m1[d] + m3[d]);
    }
    for (int c = 0; c < nc; c++) {
        int ch = base + c;
        int h[3] = {c_hx[c], c_hy[c], c_hz[c]};
        for (int d = 0; d < 3; d++) {
            if (d >= dim) {
                lo[d][ch] = 0.0;
                hi[d][ch] = 0.0;
            } else if (h[d]) {
                lo[d][ch] = m2[d];
                hi[d][ch] = m3[d];
            } else {
                lo[d][ch] = m1[d];
                hi[d][ch] = m2[d];
            }
        }
        parent[ch] = p;
        child0[ch] = -1;
        level[ch] = lvl + 1;
        count[ch] = 0;
        start[ch] = 0;
    }
    for (int j = 0; j < nch; j++) {
        chunk_panel[cb + j] = p;
        chunk_j[cb + j] = j;
    }
}

// Child of a point (tree_construction.cpp:235-269 / :77-93): strict '>'.
template <int DIM>
__device__ __forceinline__ int which_child(double x, double y, double z, double mx, double my, double mz) {
    int q = (x > mx) ? ((y > my) ? 0 : 3) : ((y > my) ? 1 : 2);
    if constexpr (DIM == 3) return (z > mz) ? q : q + 4;
    (void)z;
    (void)mz;
    return q;
}

template <int DIM>
__global__ void __launch_bounds__(TB)
    k_classify(const int *__restrict__ chunk_panel, const int *__restrict__ chunk_j, const double *__restrict__ lo0,
               const double *__restrict__ lo1, const double *__restrict__ lo2, const double *__restrict__ hi0,
               const double *__restrict__ hi1, const double *__restrict__ hi2, const int *__restrict__ count,
               const int *__restrict__ start, const double *__restrict__ sx, const double *__restrict__ sy,
               const double *__restrict__ sz, unsigned char *__restrict__ oct, int *__restrict__ hist) {
    constexpr int NC = DIM == 3 ? 8 : 4;
    __shared__ int h[NC];
    if (threadIdx.x < NC) h[threadIdx.x] = 0;
    __syncthreads();
    const int b = blockIdx.x;
    const int p = chunk_panel[b];
    const int j = chunk_j[b];
    const int rs = start[p] + j * CHUNK;
    const int re = min(start[p] + count[p], rs + CHUNK);
    const double mx = 0.5 * (lo0[p] + hi0[p]);
    const double my = 0.5 * (lo1[p] + hi1[p]);
    const double mz = DIM == 3 ? 0.5 * (lo2[p] + hi2[p]) : 0.0;
    int local[NC];
    for (int c = 0; c < NC; c++) local[c] = 0;
    for (int i = rs + threadIdx.x; i < re; i += TB) {
        int o = which_child<DIM>(sx[i], sy[i], DIM == 3 ? sz[i] : 0.0, mx, my, mz);
        oct[i] = (unsigned char)o;
        for (int c = 0; c < NC; c++) local[c] += (o == c);
    }
    for (int c = 0; c < NC; c++)
        if (local[c]) atomicAdd(&h[c], local[c]);
    __syncthreads();
    if (threadIdx.x < NC) hist[b * 8 + threadIdx.x] = h[threadIdx.x];
}

// One block per panel of the level: exclusive scan of its chunk histograms,
// child counts and child start offsets.
template <int NC>
__global__ void __launch_bounds__(TB)
    k_panel_scan(int ls, const long long *__restrict__ val, const long long *__restrict__ inc,
                 const int *__restrict__ hist, int *__restrict__ off, const int *__restrict__ child0,
                 int *__restrict__ count, int *__restrict__ start) {
    const int i = blockIdx.x;
    const long long v = val[i];
    if (!(v & 1)) return;
    const long long ex = inc[i] - v;
    const int cb = (int)(ex >> 32);
    const int nch = (int)(v >> 32);
    const int p = ls + i;
    __shared__ int sh[NC][TB];
    __shared__ int carry[NC];
    if (threadIdx.x < NC) carry[threadIdx.x] = 0;
    __syncthreads();
    for (int t0 = 0; t0 < nch; t0 += TB) {
        const int jj = t0 + threadIdx.x;
        int mine[NC];
        for (int c = 0; c < NC; c++) {
            mine[c] = jj < nch ? hist[(cb + jj) * 8 + c] : 0;
            sh[c][threadIdx.x] = mine[c];
        }
        __syncthreads();
        // Hillis-Steele inclusive scan per child
        for (int s = 1; s < TB; s <<= 1) {
            int add[NC];
            for (int c = 0; c < NC; c++) add[c] = threadIdx.x >= (unsigned)s ? sh[c][threadIdx.x - s] : 0;
            __syncthreads();
            for (int c = 0; c < NC; c++) sh[c][threadIdx.x] += add[c];
            __syncthreads();
        }
        if (jj < nch)
            for (int c = 0; c < NC; c++) off[(cb + jj) * 8 + c] = carry[c] + sh[c][threadIdx.x] - mine[c];
        __syncthreads();
        if (threadIdx.x < NC) carry[threadIdx.x] += sh[threadIdx.x][TB - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int base = child0[p];
        int s = start[p];
        for (int c = 0; c < NC; c++) {
            count[base + c] = carry[c];
            start[base + c] = s;
            s += carry[c];
        }
    }
}

// Stable scatter of each chunk's points into the child ranges.
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_scatter(const int *__restrict__ chunk_panel, const int *__restrict__ chunk_j, const int *__restrict__ count,
              const int *__restrict__ start, const int *__restrict__ child0, const int *__restrict__ off,
              const unsigned char *__restrict__ oct, const int *__restrict__ perm, const double *__restrict__ sx,
              const double *__restrict__ sy, const double *__restrict__ sz, int *__restrict__ perm2,
              double *__restrict__ sx2, double *__restrict__ sy2, double *__restrict__ sz2,
              int *__restrict__ point_leaf) {
    constexpr int NC = DIM == 3 ? 8 : 4;
    constexpr int NW = TB / kWave;
    __shared__ int wcnt[NW][NC];
    __shared__ int base[NC];
    const int b = blockIdx.x;
    const int p = chunk_panel[b];
    const int j = chunk_j[b];
    const int rs = start[p] + j * CHUNK;
    const int re = min(start[p] + count[p], rs + CHUNK);
    const int c0 = child0[p];
    if (threadIdx.x < NC) base[threadIdx.x] = start[c0 + threadIdx.x] + off[b * 8 + threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (kWave - lane));
    for (int t0 = rs; t0 < re; t0 += TB) {
        const int i = t0 + threadIdx.x;
        const bool valid = i < re;
        const int o = valid ? (int)oct[i] : -1;
        int rank = 0;
        for (int c = 0; c < NC; c++) {
            unsigned long long m = __ballot(o == c);
            if (o == c) rank = __popcll(m & lt);
            if (lane == 0) wcnt[w][c] = __popcll(m);
        }
        __syncthreads();
        if (valid) {
            int d = base[o] + rank;
            for (int ww = 0; ww < w; ww++) d += wcnt[ww][o];
            const int pi = perm[i];
            perm2[d] = pi;
            sx2[d] = sx[i];
            sy2[d] = sy[i];
            if (DIM == 3) sz2[d] = sz[i];
            point_leaf[pi] = c0 + o;  // tree_construction.cpp:273
        }
        __syncthreads();
        if (threadIdx.x < NC) {
            int s = 0;
            for (int ww = 0; ww < NW; ww++) s += wcnt[ww][threadIdx.x];
            base[threadIdx.x] += s;
        }
        __syncthreads();
    }
}

// The scatter of level L writes the refined panels' points into the buffer
// that level L+1 reads (the two buffers alternate), so a leaf's points stay
// where the level that created it wrote them: the caller-visible arrays for
// leaves of even levels, the second buffer for leaves of odd levels.  This
// pass copies the latter back, once, at the end of the build: one launch per
// odd level (panels are numbered level by level), one workgroup per panel,
// leaves only.
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_odd_leaves_back(int ls, const int *__restrict__ child0, const int *__restrict__ count,
                      const int *__restrict__ start, const int *__restrict__ perm2, const double *__restrict__ sx2,
                      const double *__restrict__ sy2, const double *__restrict__ sz2, int *__restrict__ perm,
                      double *__restrict__ sx, double *__restrict__ sy, double *__restrict__ sz) {
    const int p = ls + blockIdx.x;
    if (child0[p] >= 0) return;
    const int rs = start[p], re = rs + count[p];
    for (int i = rs + threadIdx.x; i < re; i += TB) {
        perm[i] = perm2[i];
        sx[i] = sx2[i];
        sy[i] = sy2[i];
        if (DIM == 3) sz[i] = sz2[i];
    }
}

// radius = 0.5*sqrt(x_ex*x_ex + y_ex*y_ex [+ z_ex*z_ex]) (tree_construction.cpp:296-299)
__global__ void k_radius(int dim, int pc, const double *__restrict__ lo0, const double *__restrict__ lo1,
                         const double *__restrict__ lo2, const double *__restrict__ hi0,
                         const double *__restrict__ hi1, const double *__restrict__ hi2, double *__restrict__ radius,
                         int *__restrict__ is_leaf_flag, const int *__restrict__ child0) {
    int p = blockIdx.x * TB + threadIdx.x;
    if (p >= pc) return;
    double ex = hi0[p] - lo0[p];
    double ey = hi1[p] - lo1[p];
    if (dim == 3) {
        double ez = hi2[p] - lo2[p];
        radius[p] = 0.5 * sqrt(ex * ex + ey * ey + ez * ez);
    } else {
        radius[p] = 0.5 * sqrt(ex * ex + ey * ey);
    }
    is_leaf_flag[p] = child0[p] < 0 ? 1 : 0;
}

__global__ void k_compact_leaves(int pc, const int *__restrict__ flag, const int *__restrict__ pos,
                                 int *__restrict__ leaves) {
    int p = blockIdx.x * TB + threadIdx.x;
    if (p >= pc) return;
    if (flag[p]) leaves[pos[p]] = p;
}

__global__ void k_iota(int n, int *__restrict__ perm, int *__restrict__ point_leaf) {
    int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    perm[i] = i;
    point_leaf[i] = 0;  // tree_construction.cpp:153
}

inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }

// The panel arrays of a tree, for one-launch copies.
struct PanelArrays {
    double *d[7];  // lo[3], hi[3], radius slot unused here
    int *i[5];     // count, start, parent, child0, level
};

__global__ void k_copy_panels(int n, PanelArrays src, PanelArrays dst) {
    const int p = blockIdx.x * TB + threadIdx.x;
    if (p >= n) return;
#pragma unroll
    for (int a = 0; a < 6; a++) dst.d[a][p] = src.d[a][p];
#pragma unroll
    for (int a = 0; a < 5; a++) dst.i[a][p] = src.i[a][p];
}

void grow_panels(Tree &t, int need, hipStream_t s) {
    if (need <= t.cap) return;
    int ncap = t.cap ? t.cap : 64;
    while (ncap < need) ncap *= 2;
    DBuf<double> lo[3], hi[3];
    DBuf<int> cnt(ncap), st(ncap), par(ncap), c0(ncap), lvl(ncap);
    for (int d = 0; d < 3; d++) {
        lo[d].resize(ncap);
        hi[d].resize(ncap);
    }
    // the panels already in the old arrays, in one launch (none before the
    // first allocation, when pc may already count the roots to be written)
    const int live = std::min(t.pc, t.cap);
    if (live > 0) {
        PanelArrays a{{t.lo[0].p, t.lo[1].p, t.lo[2].p, t.hi[0].p, t.hi[1].p, t.hi[2].p, nullptr},
                      {t.count.p, t.start.p, t.parent.p, t.child0.p, t.level.p}};
        PanelArrays b{{lo[0].p, lo[1].p, lo[2].p, hi[0].p, hi[1].p, hi[2].p, nullptr},
                      {cnt.p, st.p, par.p, c0.p, lvl.p}};
        k_copy_panels<<<nblk(live), TB, 0, s>>>(live, a, b);
        BTK_LAUNCH_CHECK();
    }
    for (int d = 0; d < 3; d++) {
        t.lo[d] = std::move(lo[d]);
        t.hi[d] = std::move(hi[d]);
    }
    t.count = std::move(cnt);
    t.start = std::move(st);
    t.parent = std::move(par);
    t.child0 = std::move(c0);
    t.level = std::move(lvl);
    t.cap = ncap;
}

template <int DIM>
void build_levels(Tree &t, hipStream_t s) {
    constexpr int NC = DIM == 3 ? 8 : 4;
    const int n = t.n;
    DBuf<int> perm2(n ? n : 1);
    DBuf<double> sx2(n ? n : 1), sy2(n ? n : 1), sz2(DIM == 3 && n ? n : 1);
    DBuf<unsigned char> oct(n ? n : 1);
    long long *hv = host_scratch().v;
    int lvl = 0;
    t.level_start.assign(1, 0);
    for (;;) {
        const int ls = t.level_start[lvl];
        const int np = t.pc - ls;
        DBuf<long long> val(np), inc(np);
        if (np <= kLevelFuse) {
            k_level_scan<<<1, kLevelThreads, 0, s>>>(ls, np, t.thresh, t.count.p, val.p, inc.p, hv);
            BTK_LAUNCH_CHECK();
        } else {
            k_level_flags<<<nblk(np), TB, 0, s>>>(ls, np, t.thresh, t.count.p, val.p);
            BTK_LAUNCH_CHECK();
            inclusive_sum_i64(val.p, inc.p, np, s);
            BTK_HIP(hipMemcpyAsync(hv, inc.p + (np - 1), sizeof(long long), hipMemcpyDeviceToHost, s));
        }
        BTK_HIP(hipStreamSynchronize(s));
        const long long tot = hv[0];
        const int R = (int)(tot & 0xffffffffLL);
        const int T = (int)(tot >> 32);
        if (R == 0) break;
        if (lvl + 1 >= kMaxLevels)
            throw Error("tree construction: more than " + std::to_string(kMaxLevels) +
                        " levels (more than cluster_size coincident points?)");
        const int pc_new = t.pc + NC * R;
        grow_panels(t, pc_new, s);
        DBuf<int> chunk_panel(T), chunk_j(T), hist((size_t)T * 8), off((size_t)T * 8);
        k_make_children<<<nblk(np), TB, 0, s>>>(DIM, ls, np, t.pc, lvl, val.p, inc.p, t.lo[0].p, t.lo[1].p,
                                                 t.lo[2].p, t.hi[0].p, t.hi[1].p, t.hi[2].p, t.count.p, t.start.p,
                                                 t.parent.p, t.child0.p, t.level.p, chunk_panel.p, chunk_j.p);
        BTK_LAUNCH_CHECK();
        // this level's points: the caller-visible arrays on even levels, the
        // second buffer on odd ones (k_odd_leaves_back)
        int *cp = (lvl & 1) ? perm2.p : t.perm.p, *np_ = (lvl & 1) ? t.perm.p : perm2.p;
        double *cx = (lvl & 1) ? sx2.p : t.sx.p, *nx = (lvl & 1) ? t.sx.p : sx2.p;
        double *cy = (lvl & 1) ? sy2.p : t.sy.p, *ny = (lvl & 1) ? t.sy.p : sy2.p;
        double *cz = DIM == 3 ? ((lvl & 1) ? sz2.p : t.sz.p) : t.sz.p;
        double *nz = DIM == 3 ? ((lvl & 1) ? t.sz.p : sz2.p) : sz2.p;
        k_classify<DIM><<<T, TB, 0, s>>>(chunk_panel.p, chunk_j.p, t.lo[0].p, t.lo[1].p, t.lo[2].p, t.hi[0].p,
                                          t.hi[1].p, t.hi[2].p, t.count.p, t.start.p, cx, cy, cz, oct.p, hist.p);
        BTK_LAUNCH_CHECK();
        k_panel_scan<NC><<<np, TB, 0, s>>>(ls, val.p, inc.p, hist.p, off.p, t.child0.p, t.count.p, t.start.p);
        BTK_LAUNCH_CHECK();
        k_scatter<DIM><<<T, TB, 0, s>>>(chunk_panel.p, chunk_j.p, t.count.p, t.start.p, t.child0.p, off.p, oct.p,
                                         cp, cx, cy, cz, np_, nx, ny, nz, t.point_leaf.p);
        BTK_LAUNCH_CHECK();
        t.level_start.push_back(t.pc);
        t.pc = pc_new;
        lvl++;
    }
    t.levels = lvl + 1;
    t.level_start.push_back(t.pc);
    // leaves of odd levels: their points are in the second buffer
    for (int l = 1; l < t.levels; l += 2) {
        const int ls = t.level_start[l], np = t.level_start[l + 1] - ls;
        if (np == 0) continue;
        k_odd_leaves_back<DIM><<<np, TB, 0, s>>>(ls, t.child0.p, t.count.p, t.start.p, perm2.p, sx2.p, sy2.p,
                                                  sz2.p, t.perm.p, t.sx.p, t.sy.p, t.sz.p);
        BTK_LAUNCH_CHECK();
    }
}

}  // namespace

std::unique_ptr<Tree> build_tree(int dim, int thresh, const double *x, const double *y, const double *z, int n,
                                 hipStream_t s) {
    return build_forest(dim, thresh, x, y, z, std::vector<int>{0, n}, s);
}

std::unique_ptr<Tree> build_forest(int dim, int thresh, const double *x, const double *y, const double *z,
                                   const std::vector<int> &offsets, hipStream_t s) {
    BTK_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    BTK_REQUIRE(offsets.size() >= 2 && offsets[0] == 0, "problem offsets must start at 0");
    const int nprob = (int)offsets.size() - 1;
    for (int p = 0; p < nprob; p++) BTK_REQUIRE(offsets[p + 1] >= offsets[p], "problem offsets must not decrease");
    const int n = offsets.back();
    auto t = std::make_unique<Tree>();
    t->dim = dim;
    t->n = n;
    t->thresh = thresh;
    t->nprob = nprob;
    t->pc = nprob;
    // room for a typical tree (a full tree with leaves of about thresh / 2^dim
    // points) up front; deeper trees grow the arrays level by level
    const long long guess = (2ll << dim) * (long long)n / std::max(1, thresh + 1) + 2ll * nprob + 64;
    grow_panels(*t, (int)std::min<long long>(guess, 1 << 24), s);
    const size_t nn = n ? n : 1;
    t->perm.resize(nn);
    t->point_leaf.resize(nn);
    t->sx.resize(nn);
    t->sy.resize(nn);
    t->sz.resize(dim == 3 ? nn : 1);
    if (n) {
        BTK_HIP(hipMemcpyAsync(t->sx.p, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        BTK_HIP(hipMemcpyAsync(t->sy.p, y, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        if (dim == 3) BTK_HIP(hipMemcpyAsync(t->sz.p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        k_iota<<<nblk(n), TB, 0, s>>>(n, t->perm.p, t->point_leaf.p);
        BTK_LAUNCH_CHECK();
    }
    // root boxes: chunks of at most 2048 points per problem (one workgroup
    // each, 8 points per thread), then one workgroup per problem
    std::vector<int> cb, ce, pc0(1, 0);
    const int chunk = 2048;
    for (int p = 0; p < nprob; p++) {
        int b = offsets[p];
        do {
            cb.push_back(b);
            ce.push_back(std::min(offsets[p + 1], b + chunk));
            b += chunk;
        } while (b < offsets[p + 1]);
        pc0.push_back((int)cb.size());
    }
    const int nparts = (int)cb.size();
    DBuf<int> dcb(nparts), dce(nparts), dpc0(nprob + 1), doff(nprob + 1);
    DBuf<BoxAcc> part(nparts);
    BTK_HIP(hipMemcpyAsync(dcb.p, cb.data(), sizeof(int) * nparts, hipMemcpyHostToDevice, s));
    BTK_HIP(hipMemcpyAsync(dce.p, ce.data(), sizeof(int) * nparts, hipMemcpyHostToDevice, s));
    BTK_HIP(hipMemcpyAsync(dpc0.p, pc0.data(), sizeof(int) * (nprob + 1), hipMemcpyHostToDevice, s));
    BTK_HIP(hipMemcpyAsync(doff.p, offsets.data(), sizeof(int) * (nprob + 1), hipMemcpyHostToDevice, s));
    k_box_partial<<<nparts, TB, 0, s>>>(dim, x, y, dim == 3 ? z : x, dcb.p, dce.p, part.p);
    BTK_LAUNCH_CHECK();
    k_box_final<<<nprob, TB, 0, s>>>(dim, part.p, dpc0.p, doff.p, t->lo[0].p, t->lo[1].p, t->lo[2].p, t->hi[0].p,
                                     t->hi[1].p, t->hi[2].p, t->count.p, t->start.p, t->parent.p, t->child0.p,
                                     t->level.p);
    BTK_LAUNCH_CHECK();
    BTK_HIP(hipStreamSynchronize(s));  // the host vectors above are released on return
    if (dim == 3)
        build_levels<3>(*t, s);
    else
        build_levels<2>(*t, s);

    const int pc = t->pc;
    t->radius.resize(pc);
    DBuf<int> flag(pc), pos(pc);
    k_radius<<<nblk(pc), TB, 0, s>>>(dim, pc, t->lo[0].p, t->lo[1].p, t->lo[2].p, t->hi[0].p, t->hi[1].p,
                                      t->hi[2].p, t->radius.p, flag.p, t->child0.p);
    BTK_LAUNCH_CHECK();
    exclusive_sum_i32(flag.p, pos.p, pc, s);
    long long *hv = host_scratch().v;
    int *hi = reinterpret_cast<int *>(hv);
    BTK_HIP(hipMemcpyAsync(hi, pos.p + (pc - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipMemcpyAsync(hi + 1, flag.p + (pc - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    t->n_leaves = hi[0] + hi[1];
    t->leaves.resize(t->n_leaves);
    k_compact_leaves<<<nblk(pc), TB, 0, s>>>(pc, flag.p, pos.p, t->leaves.p);
    BTK_LAUNCH_CHECK();
    t->n_leaves_deepest = t->level_start[t->levels] - t->level_start[t->levels - 1];
    return t;
}

}  // namespace btk
