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
// A work item is one wavefront: up to 64*TPT targets of one panel and a
// contiguous slice of the panel's entries.  Sources stream through a 64-entry
// LDS tile (coalesced SoA loads of the tree-sorted points, or proxy points
// generated from the source box) and every lane reads each tile entry as an
// LDS broadcast.  When a panel's entries are split over several work items,
// their partial sums meet through FP64 atomic adds; otherwise each target is
// written once.
//
// Precision.  FP64: t - s and the kernel in FP64 (the reference's arithmetic).
// Mixed (F32): t - s formed in FP64 and rounded once to FP32, the kernel in
// FP32 (hardware log2 / reciprocal / reciprocal square root), FP32 partial
// sums over one 64-source tile, added into FP64 per-target accumulators.
#pragma once

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

struct Src64 {
    double x, y, z, q;
};

__device__ __forceinline__ double sel3(int d, double a, double b, double c) { return d == 0 ? a : (d == 1 ? b : c); }

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

template <int DIM, int KERNEL, bool F32, bool EXCL>
__device__ __forceinline__ void eval_pair(const Src64 &s, const double *t, double eps2, double *acc64, float *acc32) {
    const double dx = t[0] - s.x, dy = t[1] - s.y, dz = DIM == 3 ? t[2] - s.z : 0.0;
    if constexpr (F32)
        eval_f32<DIM, KERNEL, EXCL>((float)dx, (float)dy, (float)dz, (float)s.q, (float)eps2, acc32);
    else
        eval_f64<DIM, KERNEL, EXCL>(dx, dy, dz, s.q, eps2, acc64);
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

template <int DIM, int KERNEL, bool F32, int TPT, bool FAR, bool EXCL>
__global__ void __launch_bounds__(kLanes)
    k_interact(TaskView tk, GroupView gv, TgtView tv, SrcView sv, int n1, NodeTables tabs, KParams kp,
               double *__restrict__ out, size_t comp_stride, int atomic, int same_tree) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    __shared__ Src64 tile[kLanes];
    __shared__ double cn[3][kMaxN1];
    const int lane = threadIdx.x;
    const int item = xcd_contiguous(blockIdx.x, gridDim.x);
    const int T = tk.panel[item];
    const int first = tk.first[item];
    const int e0 = tk.e0[item], e1 = tk.e1[item];
    const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
    const int ntgt = FAR ? P : tv.count[T];
    const int ts0 = FAR ? 0 : tv.start[T];
    double t[TPT][3];
    double acc[TPT][C];
#pragma unroll
    for (int k = 0; k < TPT; k++) {
        const int ti = min(first + lane + kLanes * k, ntgt - 1);
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
        for (int c = 0; c < C; c++) acc[k][c] = 0.0;
    }
    for (int e = e0; e < e1; e++) {
        const int S = gv.src[e];
        const int ty = gv.type[e];
        const bool cluster = (ty & 1) != 0;  // PC, CC: sources are the proxy points of S
        const int nsrc = cluster ? P : sv.count[S];
        if (cluster) {
            if (lane < DIM * n1) {
                const int d = lane / n1, b = lane % n1;
                cn[d][b] = proxy_node(tabs.b[b], sel3(d, sv.lo0[S], sv.lo1[S], sv.lo2[S]),
                                      sel3(d, sv.hi0[S], sv.hi1[S], sv.hi2[S]));
            }
            __syncthreads();
        }
        // With one tree for targets and sources, coincident points share a leaf,
        // so among point sources only the target leaf's own entry can hold a
        // zero-distance pair.  Proxy sources keep the test.
        const bool check = !same_tree || S == T || cluster;
        const int ss = cluster ? 0 : sv.start[S];
        const double *w = sv.W + (size_t)S * P;
        for (int t0 = 0; t0 < nsrc; t0 += kLanes) {
            const int j = t0 + lane;
            if (j < nsrc) {
                Src64 s;
                if (cluster) {
                    if constexpr (DIM == 3) {
                        const int l = j % n1, k = (j / n1) % n1, jj = j / (n1 * n1);
                        s = Src64{cn[0][jj], cn[1][k], cn[2][l], w[j]};
                    } else {
                        const int k = j % n1, jj = j / n1;
                        s = Src64{cn[0][jj], cn[1][k], 0.0, w[j]};
                    }
                } else {
                    s = Src64{sv.sx[ss + j], sv.sy[ss + j], DIM == 3 ? sv.sz[ss + j] : 0.0, sv.q[ss + j]};
                }
                tile[lane] = s;
            }
            __syncthreads();
            const int nt = min(kLanes, nsrc - t0);
            float a32[TPT][C];
#pragma unroll
            for (int k = 0; k < TPT; k++)
#pragma unroll
                for (int c = 0; c < C; c++) a32[k][c] = 0.0f;
            if (nt == kLanes && EXCL && !check) {
                // no zero-distance pair is possible here: skip the exclusion test
#pragma unroll 8
                for (int jj = 0; jj < kLanes; jj++) {
                    const Src64 s = tile[jj];
#pragma unroll
                    for (int k = 0; k < TPT; k++) eval_pair<DIM, KERNEL, F32, false>(s, t[k], kp.eps2, acc[k], a32[k]);
                }
            } else if (nt == kLanes) {
#pragma unroll 8
                for (int jj = 0; jj < kLanes; jj++) {
                    const Src64 s = tile[jj];
#pragma unroll
                    for (int k = 0; k < TPT; k++) eval_pair<DIM, KERNEL, F32, EXCL>(s, t[k], kp.eps2, acc[k], a32[k]);
                }
            } else {
                for (int jj = 0; jj < nt; jj++) {
                    const Src64 s = tile[jj];
#pragma unroll
                    for (int k = 0; k < TPT; k++) eval_pair<DIM, KERNEL, F32, EXCL>(s, t[k], kp.eps2, acc[k], a32[k]);
                }
            }
            if constexpr (F32) {
#pragma unroll
                for (int k = 0; k < TPT; k++)
#pragma unroll
                    for (int c = 0; c < C; c++) acc[k][c] += (double)a32[k][c];
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int k = 0; k < TPT; k++) {
        const int ti = first + lane + kLanes * k;
        if (ti < ntgt) {
            const size_t o = FAR ? (size_t)T * P + ti : (size_t)tv.perm[ts0 + ti];
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

// Work items of one group (near: PP+PC, far: CP+CC): blocks of `per`
// targets of a panel times slices of its entries, balanced by source count.
struct Tasks {
    int n = 0;
    DBuf<int> panel, first, e0, e1;
};
void build_tasks(const List &l, const Tree &tt, const Tree &ts, int far, int P, int per, int nsplit, Tasks &tk,
                 hipStream_t s);
// Entry slices per panel so that about 8 wavefronts per SIMD are in flight.
int choose_split(long long base_items);

// Launch of one interaction group with everything fixed but the kernel.
struct InteractLaunch {
    TaskView tk;
    GroupView gv;
    TgtView tv;
    SrcView sv;
    int n_tasks, n1, tpt, far, excl, f32, atomic;
    int same_tree;  // targets and sources are one tree (the far group never compares panels)
    NodeTables tabs;
    KParams kp;
    double *out;
    size_t comp_stride;
};

template <int DIM, int KERNEL>
void launch_interact(const InteractLaunch &L, hipStream_t s);

}  // namespace btk

namespace btk {

#define BTK_IK(F32_, TPT_, FAR_, EXCL_)                                                                          \
    k_interact<DIM, KERNEL, F32_, TPT_, FAR_, EXCL_><<<L.n_tasks, kLanes, 0, s>>>(L.tk, L.gv, L.tv, L.sv, L.n1,  \
                                                                                 L.tabs, L.kp, L.out,            \
                                                                                 L.comp_stride, L.atomic, L.same_tree)
#define BTK_IK_FE(F32_, TPT_)                                  \
    do {                                                       \
        if (L.far) {                                           \
            if (L.excl) BTK_IK(F32_, TPT_, true, true);        \
            else BTK_IK(F32_, TPT_, true, false);              \
        } else {                                               \
            if (L.excl) BTK_IK(F32_, TPT_, false, true);       \
            else BTK_IK(F32_, TPT_, false, false);             \
        }                                                      \
    } while (0)

template <int DIM, int KERNEL>
void launch_interact(const InteractLaunch &L, hipStream_t s) {
    if (L.n_tasks <= 0) return;
    if (!L.f32)
        BTK_IK_FE(false, 1);
    else if (L.tpt == 2)
        BTK_IK_FE(true, 2);
    else
        BTK_IK_FE(true, 1);
    BTK_LAUNCH_CHECK();
}

#undef BTK_IK_FE
#undef BTK_IK

extern template void launch_interact<2, BTK_KERNEL_POT>(const InteractLaunch &, hipStream_t);
extern template void launch_interact<2, BTK_KERNEL_POT_GRAD>(const InteractLaunch &, hipStream_t);
extern template void launch_interact<3, BTK_KERNEL_POT>(const InteractLaunch &, hipStream_t);
extern template void launch_interact<3, BTK_KERNEL_REF_VEL3D>(const InteractLaunch &, hipStream_t);
extern template void launch_interact<3, BTK_KERNEL_POT_GRAD>(const InteractLaunch &, hipStream_t);

}  // namespace btk
