// PP / PC / CP / CC interactions (poisson_fmm.cpp:8-26, poisson_fmm_impl.hpp:
// 10-442; poisson_vel_fmm.cpp:8-17, poisson_vel_fmm_impl.hpp:10-257).
//
// The reference runs one thread per interaction pair and adds every kernel
// evaluation to its target with an atomic.  Here the list is grouped by
// target (stably, so each target sees its sources in the reference's list
// order) and one kernel serves all four types:
//   near (FAR = false)  targets are the points of a target leaf; the entries
//                       are its PP pairs (sources: a leaf's points) and PC
//                       pairs (sources: a cluster's proxy points and weights);
//   far  (FAR = true)   targets are the proxy points of a target cluster; the
//                       entries are its CP and CC pairs.
// A work item is one workgroup of WAVES wavefronts (1, 2 or 4): up to
// 64*TPT*WAVES targets of one panel, in consecutive blocks of 64*TPT per
// wavefront, and a contiguous slice of the panel's entries.  Sources stream
// through an LDS tile of 64*WAVES entries that the whole workgroup fills
// (coalesced SoA loads of the tree-sorted points, or proxy points generated
// from the source box), so each source is read from memory once per
// workgroup; every lane reads each tile entry as an LDS broadcast.  When a
// panel's entries are split over several work items, their partial sums meet
// through FP64 atomic adds; otherwise each target is written once.
//
// Precision.  FP64: t - s and the kernel in FP64 (the reference's arithmetic).
// Mixed (F32): t - s to FP32 relative precision (near: formed in FP64 and
// rounded once, or relative to the target box centre for entries of separated
// boxes, TileMix; far: relative to the target box centre, TileRel), the
// kernel in FP32 (hardware log2 / reciprocal
// / reciprocal square root), FP32 partial sums over one 64-source tile, added
// into FP64 per-target accumulators.
#pragma once

#include <algorithm>
#include <type_traits>

#include "btk_device.hpp"

namespace btk {

constexpr int kLanes = 64;

struct SrcView {
    const int *start, *count;
    const double *sx, *sy, *sz, *q;  // sorted points and charges
    const double *lo0, *lo1, *lo2, *hi0, *hi1, *hi2;
    const double *W;  // [pc][P]
};

struct TgtView {
    const int *start, *count, *perm;
    const double *sx, *sy, *sz;
    const double *lo0, *lo1, *lo2, *hi0, *hi1, *hi2;
};

struct TaskView {
    const int *panel, *first, *e0, *e1;
};

struct GroupView {
    const int *src;
    const unsigned char *type;
};

// One source in the LDS tile.  The charge (or proxy weight) is rounded to
// FP32 once, when the tile is filled, in the mixed-precision kernels.
template <bool F32>
struct TileSrc {
    double x, y, z;
    std::conditional_t<F32, float, double> q;
};

// A source of the far (CP, CC) kernels in mixed precision: coordinates
// relative to the target cluster's box centre, rounded to FP32 once.  Far
// pairs satisfy the MAC, (r_t + r_s) / |c_s - c_t| < theta, so |t - c_t| +
// |s - c_t| is at most (1 + theta) / (1 - theta) times |t - s|, and the FP32
// rounding of the relative coordinates perturbs t - s by at most that factor
// times 2^-24 (relative): 3.4e-7 at theta = 0.7.
struct TileRel {
    float x, y, z, q;
};

__device__ __forceinline__ double sel3(int d, double a, double b, double c) { return d == 0 ? a : (d == 1 ? b : c); }

// Near pairs of separated boxes take the far field's relative coordinates
// (one FP32 subtraction per axis instead of an FP64 subtraction and a
// conversion).
// Relative to the target box centre c, |t - c| + |s - c| <= r_t + d + r_s
// (d the distance of the centres), while |t - s| >= g, the gap between the
// boxes; rounding t - c and s - c to FP32 then perturbs t - s by at most
// (r_t + d + r_s) / g times 2^-24, relative.  The pair qualifies when that
// factor is at most (1 + 0.7) / (1 - 0.7), the far field's own bound at
// theta = 0.7 (3.4e-7).
constexpr double kRelFactor = 17.0 / 3.0;
template <int DIM>
__device__ __forceinline__ bool separated_for_rel(int T, int S, const TgtView &tv, const SrcView &sv) {
    const double tl[3] = {tv.lo0[T], tv.lo1[T], DIM == 3 ? tv.lo2[T] : 0.0};
    const double th[3] = {tv.hi0[T], tv.hi1[T], DIM == 3 ? tv.hi2[T] : 0.0};
    const double sl[3] = {sv.lo0[S], sv.lo1[S], DIM == 3 ? sv.lo2[S] : 0.0};
    const double sh[3] = {sv.hi0[S], sv.hi1[S], DIM == 3 ? sv.hi2[S] : 0.0};
    double g2 = 0.0, d2 = 0.0, et2 = 0.0, es2 = 0.0;
#pragma unroll
    for (int a = 0; a < DIM; a++) {
        const double g = fmax(0.0, fmax(sl[a] - th[a], tl[a] - sh[a]));
        const double dc = 0.5 * ((sl[a] + sh[a]) - (tl[a] + th[a]));
        g2 += g * g;
        d2 += dc * dc;
        et2 += (th[a] - tl[a]) * (th[a] - tl[a]);
        es2 += (sh[a] - sl[a]) * (sh[a] - sl[a]);
    }
    const double g = sqrt(g2);
    return g > 0.0 && 0.5 * sqrt(et2) + 0.5 * sqrt(es2) + sqrt(d2) <= kRelFactor * g;
}

// A source of the near (PP, PC) kernels in mixed precision: FP64 coordinates
// (t - s formed in FP64 and rounded once: FP32 relative precision however
// close the points), or, for entries of separated boxes, FP32 coordinates
// relative to the target box centre (rx, ry, rz).
// Laid out for aligned LDS reads: the relative form (rx, ry, rz, q) in one
// 16-byte load, x and y in another; 2D drops z.
template <int DIM>
struct TileMix {
    float rx, ry, rz, q;
    double x, y;
};
template <>
struct TileMix<3> {
    float rx, ry, rz, q;
    double x, y, z, pad;
};

// Per-entry source descriptor of a grouped list: x = source panel, y = first
// sorted point of its points (0 for proxy sources), z = number of sources
// (the leaf's point count, or P proxy points), w = interact_type (bits 0-3)
// plus kRelEntry for a near entry whose boxes are separated enough for
// relative FP32 coordinates (mixed precision; separated_for_rel).
struct EntryView {
    const int4 *meta;
};
constexpr int kRelEntry = 16;


// The proxy point idx of a box: nodes cos(pi/n*j) mapped to each axis
// (bli_points_shift, bli_impl.hpp:65-77), flat index (j*(n+1)+k)*(n+1)+l.
template <int DIM>
__device__ __forceinline__ void proxy_point(int idx, int n1, const NodeTables &tabs, double lo0, double hi0,
                                            double lo1, double hi1, double lo2, double hi2, double *p) {
    if constexpr (DIM == 3) {
        const int l = idx % n1, k = (idx / n1) % n1, j = idx / (n1 * n1);
        p[0] = proxy_node(tabs.b[j], lo0, hi0);
        p[1] = proxy_node(tabs.b[k], lo1, hi1);
        p[2] = proxy_node(tabs.b[l], lo2, hi2);
    } else {
        const int k = idx % n1, j = idx / n1;
        p[0] = proxy_node(tabs.b[j], lo0, hi0);
        p[1] = proxy_node(tabs.b[k], lo1, hi1);
        p[2] = 0.0;
    }
}

// The FP64 path: t - s and the kernel in FP64.
template <int DIM, int KERNEL, bool EXCL>
__device__ __forceinline__ void eval_pair(const TileSrc<false> &s, const double *t, double eps2, double *acc64,
                                          const double2 *ltab) {
    const double dx = t[0] - s.x, dy = t[1] - s.y, dz = DIM == 3 ? t[2] - s.z : 0.0;
    eval_f64<DIM, KERNEL, EXCL, DIM == 2>(dx, dy, dz, s.q, eps2, acc64, ltab);
}

// Source j of entry m: a point of a source leaf, or a proxy point of a
// source cluster with its weight.
template <int DIM, bool F32>
__device__ __forceinline__ TileSrc<F32> load_source(const int4 &m, int j, int n1, const NodeTables &tabs,
                                                   const SrcView &sv) {
    TileSrc<F32> s{0.0, 0.0, 0.0, 0};
    if (j >= m.z) return s;
    const int S = m.x;
    if (m.w & 1) {
        double p[3];
        proxy_point<DIM>(j, n1, tabs, sv.lo0[S], sv.hi0[S], sv.lo1[S], sv.hi1[S], DIM == 3 ? sv.lo2[S] : 0.0,
                         DIM == 3 ? sv.hi2[S] : 0.0, p);
        const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
        s.x = p[0];
        s.y = p[1];
        s.z = p[2];
        s.q = (decltype(s.q))sv.W[(size_t)S * P + j];
    } else {
        const int i = m.y + j;
        s.x = sv.sx[i];
        s.y = sv.sy[i];
        if constexpr (DIM == 3) s.z = sv.sz[i];
        s.q = (decltype(s.q))sv.q[i];
    }
    return s;
}

// The same source relative to the centre c, in FP32.
template <int DIM>
__device__ __forceinline__ TileRel load_source_rel(const int4 &m, int j, int n1, const NodeTables &tabs,
                                                   const SrcView &sv, const double *c) {
    const TileSrc<true> a = load_source<DIM, true>(m, j, n1, tabs, sv);
    if (j >= m.z) return TileRel{0.0f, 0.0f, 0.0f, 0.0f};
    return TileRel{(float)(a.x - c[0]), (float)(a.y - c[1]), DIM == 3 ? (float)(a.z - c[2]) : 0.0f, a.q};
}

template <int DIM, int KERNEL, bool EXCL>
__device__ __forceinline__ void eval_pair_rel(const TileRel &s, const float *t, float eps2, float *acc32) {
    eval_f32<DIM, KERNEL, EXCL>(t[0] - s.x, t[1] - s.y, DIM == 3 ? t[2] - s.z : 0.0f, s.q, eps2, acc32);
}

// Workgroups are dispatched to the 8 XCDs round robin.  Work items are
// ordered by target panel (neighbouring leaves share most of their
// sources), so each XCD takes runs of kRun consecutive items -- its L2 then
// serves the shared source tiles -- and the runs go round robin, which keeps
// dense and sparse regions spread over all XCDs.  Items past the last full
// round keep their own index.
constexpr int kXcds = 8;
constexpr int kRun = 16;
__device__ __forceinline__ int xcd_contiguous(int bid, int n) {
    const int full = n / (kXcds * kRun) * (kXcds * kRun);
    if (bid >= full) return bid;
    const int x = bid % kXcds, r = bid / kXcds;
    return (r / kRun) * (kXcds * kRun) + x * kRun + (r % kRun);
}

// The LDS tile type of an instance.
template <int DIM, bool F32, bool FAR>
using TileOf = std::conditional_t<F32 && FAR, TileRel, std::conditional_t<F32 && !FAR, TileMix<DIM>, TileSrc<F32>>>;

// One work item, evaluated by the whole workgroup (64 * WAVES threads)
// through the LDS tile `tile` of 64 * WAVES sources.
template <int DIM, int KERNEL, bool F32, int TPT, bool FAR, bool EXCL, int WAVES, bool KAHAN = false>
__device__ __forceinline__ void interact_item(int item, const TaskView &tk, const EntryView &ev, const TgtView &tv,
                                              const SrcView &sv, int n1, const NodeTables &tabs, const KParams &kp,
                                              double *__restrict__ out, size_t comp_stride, int atomic,
                                              int same_tree, TileOf<DIM, F32, FAR> *tile,
                                              const double2 *ltab = nullptr) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    constexpr bool REL = F32 && FAR;                 // relative FP32 coordinates (TileRel)
    constexpr bool MIX = F32 && !FAR;                // FP64 or relative FP32 coordinates (TileMix)
    constexpr int TS = kLanes * WAVES;               // sources per tile
    using Tile = TileOf<DIM, F32, FAR>;
    const int tid = threadIdx.x;
    const int lane = tid % kLanes;
    const int wave = tid / kLanes;
    const int T = tk.panel[item];
    const int first = tk.first[item];
    const int e0 = tk.e0[item], e1 = tk.e1[item];
    const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
    const int ntgt = FAR ? P : tv.count[T];
    const int ts0 = FAR ? 0 : tv.start[T];
    // this wavefront's targets: first + (wave * TPT + k) * 64 + lane
    const int wfirst = first + wave * TPT * kLanes;
    const bool active = wfirst < ntgt;  // uniform over the wavefront
    double t[TPT][3];
    float tr[TPT][3];
    double acc[TPT][C];
    double ctr[3] = {0.0, 0.0, 0.0};
    if constexpr (REL || MIX) {
        ctr[0] = 0.5 * (tv.lo0[T] + tv.hi0[T]);
        ctr[1] = 0.5 * (tv.lo1[T] + tv.hi1[T]);
        if constexpr (DIM == 3) ctr[2] = 0.5 * (tv.lo2[T] + tv.hi2[T]);
    }
#pragma unroll
    for (int k = 0; k < TPT; k++) {
        const int ti = min(wfirst + lane + kLanes * k, ntgt - 1);
        if constexpr (FAR) {
            // target proxy point (poisson_fmm_impl.hpp:352-361, :416-425)
            proxy_point<DIM>(ti, n1, tabs, tv.lo0[T], tv.hi0[T], tv.lo1[T], tv.hi1[T], DIM == 3 ? tv.lo2[T] : 0.0,
                             DIM == 3 ? tv.hi2[T] : 0.0, t[k]);
        } else {
            t[k][0] = tv.sx[ts0 + ti];
            t[k][1] = tv.sy[ts0 + ti];
            t[k][2] = DIM == 3 ? tv.sz[ts0 + ti] : 0.0;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) tr[k][d] = (float)(t[k][d] - ctr[d]);
#pragma unroll
        for (int c = 0; c < C; c++) acc[k][c] = 0.0;
    }
    // near, mixed: entries of separated boxes take relative coordinates
    // (flagged when the plan built the entry descriptors)
    auto rel_entry = [&](const int4 &mm) {
        if constexpr (MIX)
            return (mm.w & kRelEntry) != 0;
        else
            return false;
    };
    auto fetch = [&](const int4 &mm, int j, bool rel) -> Tile {
        if constexpr (REL) {
            return load_source_rel<DIM>(mm, j, n1, tabs, sv, ctr);
        } else if constexpr (MIX) {
            const TileSrc<true> a = load_source<DIM, true>(mm, j, n1, tabs, sv);
            {
                Tile r{};
                r.q = a.q;
                if (rel) {
                    r.rx = (float)(a.x - ctr[0]);
                    r.ry = (float)(a.y - ctr[1]);
                    r.rz = DIM == 3 ? (float)(a.z - ctr[2]) : 0.0f;
                } else {
                    r.x = a.x;
                    r.y = a.y;
                    if constexpr (DIM == 3) r.z = a.z;
                }
                return r;
            }
        } else {
            return load_source<DIM, F32>(mm, j, n1, tabs, sv);
        }
    };
    static_assert(!KAHAN || (F32 && !FAR), "compensated sums are a variant of the mixed-precision near kernel");
    auto eval = [&](const Tile &s, int k, float *a32, float *c32, auto excl, auto relc) {
        constexpr bool X = decltype(excl)::value;
        auto f32 = [&](float dx, float dy, float dz) {
            if constexpr (KAHAN)
                eval_f32_kahan<DIM, KERNEL, X>(dx, dy, dz, s.q, (float)kp.eps2, a32, c32);
            else
                eval_f32<DIM, KERNEL, X>(dx, dy, dz, s.q, (float)kp.eps2, a32);
        };
        (void)c32;
        if constexpr (REL)
            eval_pair_rel<DIM, KERNEL, X>(s, tr[k], (float)kp.eps2, a32);
        else if constexpr (MIX) {
            if constexpr (decltype(relc)::value)
                f32(tr[k][0] - s.rx, tr[k][1] - s.ry, DIM == 3 ? tr[k][2] - s.rz : 0.0f);
            else if constexpr (DIM == 3)
                f32((float)(t[k][0] - s.x), (float)(t[k][1] - s.y), (float)(t[k][2] - s.z));
            else
                f32((float)(t[k][0] - s.x), (float)(t[k][1] - s.y), 0.0f);
        } else {
            static_assert(!F32, "mixed precision takes TileRel (far) or TileMix (near)");
            eval_pair<DIM, KERNEL, X>(s, t[k], kp.eps2, acc[k], ltab);
        }
    };
    // The entries' sources stream through the tile TS at a time, as one
    // sequence over all entries of the item; the next tile's loads are
    // issued before the current tile is evaluated.
    int e = e0, j0 = 0;
    int4 m = make_int4(0, 0, 0, 0);
    while (e < e1 && (m = ev.meta[e]).z == 0) e++;
    Tile next{};
    bool rel = false;
    if (e < e1) {
        rel = rel_entry(m);
        next = fetch(m, tid, rel);
    }
    while (e < e1) {
        tile[tid] = next;
        __syncthreads();
        const int nt = min(TS, m.z - j0);
        // With one tree for targets and sources, coincident points share a
        // leaf, so among point sources only the target leaf's own entry can
        // hold a zero-distance pair.  Proxy sources keep the test, except
        // those of separated boxes.
        const bool check = (!same_tree || m.x == T || (m.w & 1)) && !rel;
        int en = e, jn = j0 + TS;
        int4 mn = m;
        bool reln = rel;
        if (jn >= m.z) {
            jn = 0;
            en++;
            while (en < e1 && (mn = ev.meta[en]).z == 0) en++;
            if (en < e1) reln = rel_entry(mn);
        }
        if (en < e1) next = fetch(mn, jn + tid, reln);
        // The tile in chunks of 64 sources: each chunk's FP32 partial sums
        // go into the FP64 accumulators, so the FP32 blocking stays at 64
        // terms whatever the tile size.
        auto run_chunk = [&](int c0, int cn, auto excl, auto relc) {
            float a32[TPT][C], c32[TPT][C];
#pragma unroll
            for (int k = 0; k < TPT; k++)
#pragma unroll
                for (int c = 0; c < C; c++) a32[k][c] = c32[k][c] = 0.0f;
            if (cn == kLanes) {
#pragma unroll 8
                for (int jj = 0; jj < kLanes; jj++) {
                    const Tile s = tile[c0 + jj];
#pragma unroll
                    for (int k = 0; k < TPT; k++) eval(s, k, a32[k], c32[k], excl, relc);
                }
            } else {
                for (int jj = 0; jj < cn; jj++) {
                    const Tile s = tile[c0 + jj];
#pragma unroll
                    for (int k = 0; k < TPT; k++) eval(s, k, a32[k], c32[k], excl, relc);
                }
            }
            if constexpr (F32) {
#pragma unroll
                for (int k = 0; k < TPT; k++)
#pragma unroll
                    for (int c = 0; c < C; c++)
                        acc[k][c] += KAHAN ? (double)a32[k][c] - (double)c32[k][c] : (double)a32[k][c];
            }
        };
        auto run_tile = [&](auto relc) {
            for (int c0 = 0; c0 < nt; c0 += kLanes) {
                const int cn = min(kLanes, nt - c0);
                if (EXCL && !check)  // no zero-distance pair is possible here: no exclusion test
                    run_chunk(c0, cn, std::false_type{}, relc);
                else
                    run_chunk(c0, cn, std::bool_constant<EXCL>{}, relc);
            }
        };
        if (active) {  // a wavefront without targets only helps fill the tiles
            if constexpr (MIX) {
                if (rel)
                    run_tile(std::true_type{});
                else
                    run_tile(std::false_type{});
            } else {
                run_tile(std::false_type{});
            }
        }
        __syncthreads();
        e = en;
        j0 = jn;
        m = mn;
        rel = reln;
    }
#pragma unroll
    for (int k = 0; k < TPT; k++) {
        const int ti = wfirst + lane + kLanes * k;
        if (ti < ntgt) {
            const size_t o = FAR ? (size_t)T * P + ti : (size_t)(tv.perm ? tv.perm[ts0 + ti] : ts0 + ti);
#pragma unroll
            for (int c = 0; c < C; c++) {
                double scale = comp_scale<DIM, KERNEL>(c, kp.gfc);
                if constexpr (F32) scale = comp_scale_f32<DIM, KERNEL>(c, kp.gfc);
                const double v = scale * acc[k][c];
                double *dst = out + c * comp_stride + o;
                if (atomic)
                    unsafeAtomicAdd(dst, v);
                else
                    *dst += v;
            }
        }
    }
}

// The mixed-precision kernels ask for at least 6 wavefronts per SIMD (at most
// 80 VGPRs); left to itself the compiler spends up to 113 on unrolling and
// runs 4.  The FP64 kernels keep the compiler's choice.
template <int DIM, int KERNEL, bool F32, int TPT, bool FAR, bool EXCL, int WAVES, bool KAHAN = false>
__global__ void __launch_bounds__(kLanes * WAVES, F32 ? 6 : 0)
    k_interact(TaskView tk, EntryView ev, TgtView tv, SrcView sv, int n1, NodeTables tabs, KParams kp,
               double *__restrict__ out, size_t comp_stride, int atomic, int same_tree) {
    __shared__ TileOf<DIM, F32, FAR> tile[kLanes * WAVES];
    constexpr bool LOGTAB = DIM == 2 && !F32;  // FP64 2D: the table-driven log
    __shared__ double2 ltab[LOGTAB ? kLogTab : 1];
    if constexpr (LOGTAB) {
        fill_log_tab(ltab, threadIdx.x, blockDim.x);
        __syncthreads();
    }
    interact_item<DIM, KERNEL, F32, TPT, FAR, EXCL, WAVES, KAHAN>(xcd_contiguous(blockIdx.x, gridDim.x), tk, ev, tv,
                                                                  sv, n1, tabs, kp, out, comp_stride, atomic,
                                                                  same_tree, tile, LOGTAB ? ltab : nullptr);
}

// One group (near or far, one degree) of a work queue.  Its source proxy
// weights are W + woff, its outputs out + ooff (near) or Psi + ooff (far).
struct QueueGroup {
    TaskView tk;
    size_t woff, ooff, comp_stride;
    int n1, far, atomic, first;  // first: the group's first item in the flattened numbering
    NodeTables tabs;
};

// Persistent interactions: a grid sized to the GPU's resident workgroups
// drains one queue of work items of every near and far group (biggest
// first), taking the next item by atomic ticket.  *ticket must be zero at
// launch; every workgroup leaves once the ticket passes n.
template <int DIM, int KERNEL, bool F32, int NTPT, int FTPT, bool EXCL>
__global__ void __launch_bounds__(kLanes, F32 ? 6 : 0)
    k_interact_queue(const QueueGroup *__restrict__ groups, const int2 *__restrict__ queue, int n,
                     int *__restrict__ ticket, EntryView ev, TgtView tv, SrcView sv, KParams kp, int same_tree,
                     const double *__restrict__ W, double *__restrict__ out, double *__restrict__ Psi) {
    __shared__ TileOf<DIM, F32, false> tn[kLanes];
    __shared__ TileOf<DIM, F32, true> tf[kLanes];
    __shared__ int s_item;
    constexpr bool LOGTAB = DIM == 2 && !F32;  // FP64 2D: the table-driven log
    __shared__ double2 ltab[LOGTAB ? kLogTab : 1];
    if constexpr (LOGTAB) fill_log_tab(ltab, threadIdx.x, blockDim.x);  // the loop's first barrier orders it
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(ticket, 1);
        __syncthreads();
        const int q = s_item;
        __syncthreads();  // every lane has read s_item before the next ticket overwrites it
        if (q >= n) break;
        const int2 gi = queue[q];
        const QueueGroup &g = groups[gi.x];
        SrcView s2 = sv;
        s2.W = W + g.woff;
        if (g.far)
            interact_item<DIM, KERNEL, F32, FTPT, true, EXCL, 1>(gi.y, g.tk, ev, tv, s2, g.n1, g.tabs, kp, Psi + g.ooff,
                                                                 g.comp_stride, g.atomic, 0, tf,
                                                                 LOGTAB ? ltab : nullptr);
        else
            interact_item<DIM, KERNEL, F32, NTPT, false, EXCL, 1>(gi.y, g.tk, ev, tv, s2, g.n1, g.tabs, kp,
                                                                  out + g.ooff, g.comp_stride, g.atomic, same_tree,
                                                                  tn, LOGTAB ? ltab : nullptr);
    }
}

// Work items of one entry group of the list (near: PP+PC, far: CP+CC, at one
// degree): blocks of `per` targets of a panel times slices of its entries,
// balanced by source count.
struct Tasks {
    int n = 0;
    DBuf<int> panel, first, e0, e1;
};
// With meta, entries whose source count meta[e].z is zero are left out of the
// balance (another kernel evaluates them).
void build_tasks(const List &l, const Tree &tt, const Tree &ts, int grp, int P, int per, int nsplit, Tasks &tk,
                 hipStream_t s, const int4 *meta = nullptr);
// Entry slices per panel so that the launch has about target_items work
// items (at most 16 slices per panel).
int choose_split(long long base_items, long long target_items);

// Mutual PP pairs of one tree (btk_mutual.hip): workgroups of `waves`
// wavefronts, each item a block of targets of a leaf T and a slice of T's
// owned mutual entries (mmeta: source descriptors, zero count if not owned).
struct Mutual {
    int n = 0;
    int tpt = 1, waves = 1;
    DBuf<int4> mmeta;
    Tasks tasks;
};
// Finds the mutual pairs of a one-tree list, sets the source count of both
// their entries in meta to zero and builds the work items.
void build_mutual(const List &l, const Tree &t, int tpt, int4 *meta, Mutual &mu, hipStream_t s);
// Mixed-precision evaluation of the mutual pairs, both directions, into out[C][n]
// (in the tree's sorted order when sorted_out, else in the original order).
template <int DIM, int KERNEL>
void launch_mutual(const Mutual &mu, const Tree &t, const double *q_sorted, const KParams &kp, double *out,
                   bool sorted_out, hipStream_t s);
extern template void launch_mutual<2, BTK_KERNEL_POT>(const Mutual &, const Tree &, const double *, const KParams &,
                                                      double *, bool, hipStream_t);
extern template void launch_mutual<2, BTK_KERNEL_POT_GRAD>(const Mutual &, const Tree &, const double *,
                                                           const KParams &, double *, bool, hipStream_t);
extern template void launch_mutual<3, BTK_KERNEL_POT>(const Mutual &, const Tree &, const double *, const KParams &,
                                                      double *, bool, hipStream_t);
extern template void launch_mutual<3, BTK_KERNEL_REF_VEL3D>(const Mutual &, const Tree &, const double *,
                                                            const KParams &, double *, bool, hipStream_t);
extern template void launch_mutual<3, BTK_KERNEL_POT_GRAD>(const Mutual &, const Tree &, const double *,
                                                           const KParams &, double *, bool, hipStream_t);

// CC pairs sharing a kernel matrix (btk_mrhs.hip), mixed precision, 3D.
// Panel boxes are exact halves of their parents, so the P x P kernel matrix
// of a CC pair depends only on the problem (a forest's problems have roots of
// their own), the two levels and the offset of the box centres (a multiple of
// half the finer box per axis when targets and sources share the root box;
// keys are checked against the boxes' extents).  The CC entries of one far
// group are keyed by (problem, target level, source level, offset) and every key held
// by at least min_group entries becomes one MFMA product: the matrix (from
// the key's first pair) times the pairs' proxy weights as columns.  A work
// item (x = first entry of the key in ent_T/ent_S, y = entries, z = first
// target proxy (0), w = first entry of the item) covers up to 32 pairs
// against all target proxies, 128 at a time.  Build sets the source count of the entries it takes to zero
// in meta, so the regular far kernel skips them.
struct CCGemm {
    int n_items = 0;
    int nv = 0;            // keyed entries (ent_T, ent_S)
    long long shared = 0;  // entries taken
    DBuf<int4> items;
    DBuf<int> ent_T, ent_S;  // target and source panel of each keyed entry, in key order
    DBuf<int> tgt_off, tgt_k;  // per target panel, its keyed entries that items cover (CSR)
};
constexpr int kCcMinGroup = 4;
void build_ccgemm(const List &l, const Tree &tt, const Tree &ts, int grp, int min_group, int4 *meta,
                  CCGemm &cg, hipStream_t s);
// Split-FP16 source weights of one degree (the products' A operand): rows
// of cc_row_stride(P) halves per source panel, hi and lo parts, and each
// panel's power-of-two scale, as written by weights_to_f16.
inline int cc_row_stride(int P) { return (P + 63) / 64 * 64; }
// The products of each keyed entry are staged in FP32, C rows of
// cc_stage_stride(P) floats, and added into Psi per target panel.
inline int cc_stage_stride(int P) { return (P + 31) / 32 * 32; }
void launch_ccgemm(const CCGemm &cg, const Tree &tt, const Tree &ts, const unsigned short *Wh,
                   const unsigned short *Wl, const int *wexp, int n1, const NodeTables &tabs, const KParams &kp,
                   int kernel, float *stage, double *Psi, size_t comp_stride, hipStream_t s);
void weights_to_f16(int pc, int P, int Ps, const double *W, unsigned short *Wh, unsigned short *Wl, int *wexp,
                    hipStream_t s);

// Everything the interaction launches of one list need besides the charges,
// the weights and the outputs: entry descriptors and the work items of the
// near (PP+PC) and far (CP+CC) groups.  Built once (allocations, host
// reads); launch_interactions then only launches kernels, so it can be
// captured in a HIP graph.
struct InteractPlan {
    btk_run_config cfg{};
    btk_options o{};
    const Tree *tt = nullptr, *ts = nullptr;
    KParams kp{};
    int ndeg = 1;
    int deg[4] = {0, 0, 0, 0};
    NodeTables tabs[4]{};  // per degree index
    size_t woff[4]{}, poff[4]{};  // each degree's block in W and Psi (proxy_offset)
    DBuf<int4> meta;
    struct Group {
        bool present = false;
        int tpt = 1, waves = 1, nsplit = 1;
        bool mutual = false;
        Tasks tk;
        Mutual mu;
        CCGemm cc;
    } g[2][4];  // [far][degree index]
    long long cc_shared = 0;  // CC entries evaluated through shared kernel matrices
    bool sorted_out = false;  // near output in the target tree's sorted order
    // split-FP16 source weights of the degrees with CC products (weights_to_f16):
    // hi and lo rows, each panel's scale, and each degree's offsets in them
    DBuf<unsigned short> wh, wl;
    DBuf<int> wexp;
    size_t w16off[4]{};
    DBuf<float> ccstage;  // the products of one degree's keyed entries (launches of the degrees run in order)
    // btk_options.schedule = BTK_SCHEDULE_QUEUE: the work items of every
    // group as one queue, biggest first, drained by one persistent launch
    struct Queue {
        int n = 0, ntpt = 1, ftpt = 1;
        DBuf<QueueGroup> groups;
        DBuf<int2> order;
        DBuf<int> ticket;
    } queue;
};

// Launch of one interaction group with everything fixed but the kernel.
struct InteractLaunch {
    TaskView tk;
    EntryView ev;
    TgtView tv;
    SrcView sv;
    int n_tasks, n1, tpt, waves, far, excl, f32, atomic;
    int kahan;  // mixed-precision near launch: Kahan-compensated FP32 sums (one target per lane)
    int same_tree;  // targets and sources are one tree (the far group never compares panels)
    NodeTables tabs;
    KParams kp;
    double *out;
    size_t comp_stride;
};

template <int DIM, int KERNEL>
void launch_interact(const InteractLaunch &L, hipStream_t s);

// The persistent launch of a plan's queue (targets, sources and the kernel
// parameters as for the group launches).
struct QueueLaunch {
    const InteractPlan::Queue *q;
    EntryView ev;
    TgtView tv;
    SrcView sv;
    KParams kp;
    int same_tree, excl, f32;
    const double *W;
    double *out, *Psi;
};
template <int DIM, int KERNEL>
void launch_queue(const QueueLaunch &L, hipStream_t s);

}  // namespace btk

namespace btk {

#define BTK_IK(F32_, TPT_, FAR_, EXCL_, W_)                                                                 \
    k_interact<DIM, KERNEL, F32_, TPT_, FAR_, EXCL_, W_><<<L.n_tasks, kLanes * W_, 0, s>>>(                  \
        L.tk, L.ev, L.tv, L.sv, L.n1, L.tabs, L.kp, L.out, L.comp_stride, L.atomic, L.same_tree)
#define BTK_IK_FE(F32_, TPT_, W_)                                  \
    do {                                                           \
        if (L.far) {                                               \
            if (L.excl) BTK_IK(F32_, TPT_, true, true, W_);        \
            else BTK_IK(F32_, TPT_, true, false, W_);              \
        } else {                                                   \
            if (L.excl) BTK_IK(F32_, TPT_, false, true, W_);       \
            else BTK_IK(F32_, TPT_, false, false, W_);             \
        }                                                          \
    } while (0)
#define BTK_IK_W(F32_, TPT_)                            \
    do {                                                \
        if (L.waves == 4) BTK_IK_FE(F32_, TPT_, 4);     \
        else if (L.waves == 2) BTK_IK_FE(F32_, TPT_, 2); \
        else BTK_IK_FE(F32_, TPT_, 1);                  \
    } while (0)

#define BTK_IKK(EXCL_, W_)                                                                                  \
    k_interact<DIM, KERNEL, true, 1, false, EXCL_, W_, true><<<L.n_tasks, kLanes * W_, 0, s>>>(              \
        L.tk, L.ev, L.tv, L.sv, L.n1, L.tabs, L.kp, L.out, L.comp_stride, L.atomic, L.same_tree)
#define BTK_IKK_W(EXCL_)                                \
    do {                                                \
        if (L.waves == 4) BTK_IKK(EXCL_, 4);            \
        else if (L.waves == 2) BTK_IKK(EXCL_, 2);       \
        else BTK_IKK(EXCL_, 1);                         \
    } while (0)

template <int DIM, int KERNEL>
void launch_interact(const InteractLaunch &L, hipStream_t s) {
    if (L.n_tasks <= 0) return;
    if (L.kahan && L.f32 && !L.far) {
        if (L.excl)
            BTK_IKK_W(true);
        else
            BTK_IKK_W(false);
    } else if (!L.f32)
        BTK_IK_W(false, 1);
    else if (L.tpt == 2)
        BTK_IK_W(true, 2);
    else
        BTK_IK_W(true, 1);
    BTK_LAUNCH_CHECK();
}

#undef BTK_IKK_W
#undef BTK_IKK
#undef BTK_IK_W
#undef BTK_IK_FE
#undef BTK_IK

namespace {
__global__ void k_ticket_reset(int *p) { *p = 0; }
}  // namespace
int queue_grid(const void *kernel);

#define BTK_IQ(F32_, NT_, FT_, EXCL_)                                                                          \
    do {                                                                                                      \
        auto kf = k_interact_queue<DIM, KERNEL, F32_, NT_, FT_, EXCL_>;                                        \
        const int g = std::min(L.q->n, queue_grid(reinterpret_cast<const void *>(kf)));                       \
        kf<<<g, kLanes, 0, s>>>(L.q->groups.p, L.q->order.p, L.q->n, L.q->ticket.p, L.ev, L.tv, L.sv, L.kp,    \
                                L.same_tree, L.W, L.out, L.Psi);                                              \
    } while (0)
#define BTK_IQ_E(F32_, NT_, FT_)                    \
    do {                                            \
        if (L.excl) BTK_IQ(F32_, NT_, FT_, true);   \
        else BTK_IQ(F32_, NT_, FT_, false);         \
    } while (0)

template <int DIM, int KERNEL>
void launch_queue(const QueueLaunch &L, hipStream_t s) {
    if (L.q->n <= 0) return;
    k_ticket_reset<<<1, 1, 0, s>>>(L.q->ticket.p);  // a kernel, not a memset: the launch may be captured in a graph
    BTK_LAUNCH_CHECK();
    const int nt = L.q->ntpt, ft = L.q->ftpt;
    if (!L.f32)
        BTK_IQ_E(false, 1, 1);
    else if (nt == 1 && ft == 1)
        BTK_IQ_E(true, 1, 1);
    else if (nt == 1)
        BTK_IQ_E(true, 1, 2);
    else if (ft == 1)
        BTK_IQ_E(true, 2, 1);
    else
        BTK_IQ_E(true, 2, 2);
    BTK_LAUNCH_CHECK();
}

#undef BTK_IQ_E
#undef BTK_IQ

extern template void launch_queue<2, BTK_KERNEL_POT>(const QueueLaunch &, hipStream_t);
extern template void launch_queue<2, BTK_KERNEL_POT_GRAD>(const QueueLaunch &, hipStream_t);
extern template void launch_queue<3, BTK_KERNEL_POT>(const QueueLaunch &, hipStream_t);
extern template void launch_queue<3, BTK_KERNEL_REF_VEL3D>(const QueueLaunch &, hipStream_t);
extern template void launch_queue<3, BTK_KERNEL_POT_GRAD>(const QueueLaunch &, hipStream_t);
extern template void launch_interact<2, BTK_KERNEL_POT>(const InteractLaunch &, hipStream_t);
extern template void launch_interact<2, BTK_KERNEL_POT_GRAD>(const InteractLaunch &, hipStream_t);
extern template void launch_interact<3, BTK_KERNEL_POT>(const InteractLaunch &, hipStream_t);
extern template void launch_interact<3, BTK_KERNEL_REF_VEL3D>(const InteractLaunch &, hipStream_t);
extern template void launch_interact<3, BTK_KERNEL_POT_GRAD>(const InteractLaunch &, hipStream_t);

}  // namespace btk
