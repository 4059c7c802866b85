// PP / PC / CP / CC interactions (poisson_fmm.cpp:8-26, poisson_fmm_impl.hpp:
// 10-442; poisson_vel_fmm.cpp:8-17, poisson_vel_fmm_impl.hpp:10-257) and the
// FP64 direct sum (direct_sum_impl.hpp:9-94).
//
// The reference runs one thread per interaction pair and adds every kernel
// evaluation to its target with an atomic.  Here the list is grouped by
// target (stably, so each target sees its sources in the reference's list
// order) and every target is owned by one lane:
//   near kernel  one wavefront owns up to 64*TPT points of one target leaf
//                and streams the sources of the leaf's PP pairs (the source
//                leaf's points) and PC pairs (the source cluster's proxy
//                points and weights) through 64-entry LDS tiles;
//   far kernel   a 256-lane workgroup owns a block of the proxy points of one
//                target cluster and streams the sources of its CP pairs
//                (points) and CC pairs (proxy points); when the cluster has
//                fewer proxy points than lanes, the waves split each source
//                tile and their partial sums meet in LDS.
// Source tiles are coalesced SoA loads of the tree-sorted coordinates, read
// by every lane as LDS broadcasts.  Results are written once per target.
//
// Precision.  The FP64 path evaluates (t - s) on absolute coordinates as the
// reference does.  The mixed path (BTK_PREC_MIXED) forms t - s in FP64,
// rounds it once to FP32 and evaluates the kernel in FP32 with the hardware
// log2 / reciprocal / reciprocal square root; FP32 partial sums run over at
// most one source tile and are then added into FP64 per-target accumulators.
#include "btk_device.hpp"

namespace btk {
namespace {

constexpr int TB = 256;
constexpr int NEAR_LANES = 64;
inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }

struct SrcView {
    const int *start, *count;
    const double *sx, *sy, *sz, *q;  // sorted points and charges
    const double *lo0, *lo1, *lo2, *hi0, *hi1, *hi2;
    const double *W;  // [pc][P]
};

struct Src64 {
    double x, y, z, q;
};

template <bool F32>
using TileT = Src64;  // FP64 coordinates in both paths: differences are formed in FP64

template <bool F32>
__device__ __forceinline__ Src64 make_src(double x, double y, double z, double q, double, double, double) {
    return Src64{x, y, z, q};
}

template <int DIM>
__device__ __forceinline__ double box_lo(const SrcView &sv, int d, int S) {
    return d == 0 ? sv.lo0[S] : (d == 1 ? sv.lo1[S] : sv.lo2[S]);
}
template <int DIM>
__device__ __forceinline__ double box_hi(const SrcView &sv, int d, int S) {
    return d == 0 ? sv.hi0[S] : (d == 1 ? sv.hi1[S] : sv.hi2[S]);
}

// Source j of a list entry: point j of leaf S, or proxy point j of cluster S
// (nodes cn from bli_points_shift, poisson_fmm_impl.hpp:292-294) with weight W.
template <int DIM, bool F32>
__device__ __forceinline__ TileT<F32> load_src(const SrcView &sv, bool cluster, int S, int j, int n1, int P,
                                               const double (*cn)[kMaxN1], double cx, double cy, double cz) {
    if (!cluster) {
        const int i = sv.start[S] + j;
        return make_src<F32>(sv.sx[i], sv.sy[i], DIM == 3 ? sv.sz[i] : 0.0, sv.q[i], cx, cy, cz);
    }
    double x, y, z = 0.0;
    if constexpr (DIM == 3) {
        const int l = j % n1, k = (j / n1) % n1, jj = j / (n1 * n1);
        x = cn[0][jj];
        y = cn[1][k];
        z = cn[2][l];
    } else {
        const int k = j % n1, jj = j / n1;
        x = cn[0][jj];
        y = cn[1][k];
    }
    return make_src<F32>(x, y, z, sv.W[(size_t)S * P + j], cx, cy, cz);
}

template <int DIM, int KERNEL, bool F32>
__device__ __forceinline__ void eval_any(const TileT<F32> &s, const double *t64, const float *t32, double eps2,
                                         int excl, double *acc64, float *acc32) {
    if constexpr (F32) {
        (void)t32;
        eval_f32<DIM, KERNEL>((float)(t64[0] - s.x), (float)(t64[1] - s.y), DIM == 3 ? (float)(t64[2] - s.z) : 0.0f,
                              (float)s.q, (float)eps2, excl, acc32);
    } else {
        eval_f64<DIM, KERNEL>(t64[0], t64[1], t64[2], s.x, s.y, s.z, s.q, eps2, excl, acc64);
    }
}

template <int DIM, int KERNEL>
__device__ __forceinline__ void finish_write(const double *acc, double gfc, double *out, size_t stride, size_t at) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
#pragma unroll
    for (int c = 0; c < C; c++) out[c * stride + at] += comp_scale<DIM, KERNEL>(c, gfc) * acc[c];
}

template <int DIM, int KERNEL, bool F32>
__device__ __forceinline__ double scale_of(int c, double gfc) {
    if constexpr (F32)
        return comp_scale_f32<DIM, KERNEL>(c, gfc);
    else
        return comp_scale<DIM, KERNEL>(c, gfc);
}

// ---------------------------------------------------------------------------
// PP + PC: one wavefront per (target leaf, block of 64*TPT targets).
// ---------------------------------------------------------------------------
template <int DIM, int KERNEL, bool F32, int TPT>
__global__ void __launch_bounds__(NEAR_LANES)
    k_near(const int *__restrict__ task_panel, const int *__restrict__ task_first, const int *__restrict__ goff,
           const int *__restrict__ gsrc, const unsigned char *__restrict__ gtype, const int *__restrict__ tstart,
           const int *__restrict__ tcount, const double *__restrict__ tlo0, const double *__restrict__ tlo1,
           const double *__restrict__ tlo2, const double *__restrict__ thi0, const double *__restrict__ thi1,
           const double *__restrict__ thi2, const double *__restrict__ tx_, const double *__restrict__ ty_,
           const double *__restrict__ tz_, const int *__restrict__ tperm, SrcView sv, int n1, NodeTables tabs,
           KParams kp, double *__restrict__ out, int M) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    __shared__ TileT<F32> tile[NEAR_LANES];
    __shared__ double cn[3][kMaxN1];
    const int lane = threadIdx.x;
    const int T = task_panel[blockIdx.x];
    const int first = task_first[blockIdx.x];
    const int cnt = tcount[T], ts0 = tstart[T];
    const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
    // frame centre: the target leaf's box centre (mixed path only)
    const double cx = F32 ? 0.5 * (tlo0[T] + thi0[T]) : 0.0;
    const double cy = F32 ? 0.5 * (tlo1[T] + thi1[T]) : 0.0;
    const double cz = (F32 && DIM == 3) ? 0.5 * (tlo2[T] + thi2[T]) : 0.0;
    double t64[TPT][3];
    float t32[TPT][3];
    double acc[TPT][C];
#pragma unroll
    for (int k = 0; k < TPT; k++) {
        const int ti = min(first + lane + NEAR_LANES * k, cnt - 1);
        const int i = ts0 + ti;
        t64[k][0] = tx_[i];
        t64[k][1] = ty_[i];
        t64[k][2] = DIM == 3 ? tz_[i] : 0.0;
        t32[k][0] = (float)(t64[k][0] - cx);
        t32[k][1] = (float)(t64[k][1] - cy);
        t32[k][2] = (float)(t64[k][2] - cz);
#pragma unroll
        for (int c = 0; c < C; c++) acc[k][c] = 0.0;
    }
    const int e0 = goff[T], e1 = goff[T + 1];
    for (int e = e0; e < e1; e++) {
        const int S = gsrc[e];
        const bool cluster = gtype[e] == 1;
        const int nsrc = cluster ? P : sv.count[S];
        if (cluster) {
            if (lane < DIM * n1) {
                const int d = lane / n1, b = lane % n1;
                cn[d][b] = proxy_node(tabs.b[b], box_lo<DIM>(sv, d, S), box_hi<DIM>(sv, d, S));
            }
            __syncthreads();
        }
        for (int t0 = 0; t0 < nsrc; t0 += NEAR_LANES) {
            const int j = t0 + lane;
            if (j < nsrc) tile[lane] = load_src<DIM, F32>(sv, cluster, S, j, n1, P, cn, cx, cy, cz);
            __syncthreads();
            const int nt = min(NEAR_LANES, nsrc - t0);
            float a32[TPT][C > 0 ? C : 1];
#pragma unroll
            for (int k = 0; k < TPT; k++)
#pragma unroll
                for (int c = 0; c < C; c++) a32[k][c] = 0.0f;
            if (nt == NEAR_LANES) {
#pragma unroll 8
                for (int jj = 0; jj < NEAR_LANES; jj++) {
                    const TileT<F32> s = tile[jj];
#pragma unroll
                    for (int k = 0; k < TPT; k++)
                        eval_any<DIM, KERNEL, F32>(s, t64[k], t32[k], kp.eps2, kp.excl, acc[k], a32[k]);
                }
            } else {
                for (int jj = 0; jj < nt; jj++) {
                    const TileT<F32> s = tile[jj];
#pragma unroll
                    for (int k = 0; k < TPT; k++)
                        eval_any<DIM, KERNEL, F32>(s, t64[k], t32[k], kp.eps2, kp.excl, acc[k], a32[k]);
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
        const int ti = first + lane + NEAR_LANES * k;
        if (ti < cnt) {
            const size_t o = (size_t)tperm[ts0 + ti];
#pragma unroll
            for (int c = 0; c < C; c++) out[c * (size_t)M + o] += scale_of<DIM, KERNEL, F32>(c, kp.gfc) * acc[k][c];
        }
    }
}

// ---------------------------------------------------------------------------
// CP + CC: one 256-lane workgroup per (target cluster, block of TGT proxy
// points); SPLIT = 256 / TGT groups of lanes share each source tile.
// ---------------------------------------------------------------------------
template <int DIM, int KERNEL, bool F32, int TGT>
__global__ void __launch_bounds__(TB)
    k_far(const int *__restrict__ task_panel, const int *__restrict__ task_first, const int *__restrict__ goff,
          const int *__restrict__ gsrc, const unsigned char *__restrict__ gtype, const double *__restrict__ tlo0,
          const double *__restrict__ tlo1, const double *__restrict__ tlo2, const double *__restrict__ thi0,
          const double *__restrict__ thi1, const double *__restrict__ thi2, SrcView sv, int n1, NodeTables tabs,
          KParams kp, double *__restrict__ Psi, int pct) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    constexpr int SPLIT = TB / TGT;
    __shared__ TileT<F32> tile[TB];
    __shared__ double cn[3][kMaxN1];
    __shared__ double red[SPLIT > 1 ? SPLIT : 1][TGT][C];
    const int slot = threadIdx.x % TGT;
    const int g = threadIdx.x / TGT;
    const int T = task_panel[blockIdx.x];
    const int first = task_first[blockIdx.x];
    const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
    const int idx = first + slot;
    const int id = idx < P ? idx : P - 1;
    const double cx = F32 ? 0.5 * (tlo0[T] + thi0[T]) : 0.0;
    const double cy = F32 ? 0.5 * (tlo1[T] + thi1[T]) : 0.0;
    const double cz = (F32 && DIM == 3) ? 0.5 * (tlo2[T] + thi2[T]) : 0.0;
    // target proxy point (poisson_fmm_impl.hpp:352-361)
    double t64[3] = {0.0, 0.0, 0.0};
    if (DIM == 3) {
        const int l = id % n1, k = (id / n1) % n1, j = id / (n1 * n1);
        t64[0] = proxy_node(tabs.b[j], tlo0[T], thi0[T]);
        t64[1] = proxy_node(tabs.b[k], tlo1[T], thi1[T]);
        t64[2] = proxy_node(tabs.b[l], tlo2[T], thi2[T]);
    } else {
        const int k = id % n1, j = id / n1;
        t64[0] = proxy_node(tabs.b[j], tlo0[T], thi0[T]);
        t64[1] = proxy_node(tabs.b[k], tlo1[T], thi1[T]);
    }
    const float t32[3] = {(float)(t64[0] - cx), (float)(t64[1] - cy), (float)(t64[2] - cz)};
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.0;
    const int e0 = goff[pct + T], e1 = goff[pct + T + 1];
    for (int e = e0; e < e1; e++) {
        const int S = gsrc[e];
        const bool cluster = gtype[e] == 3;
        const int nsrc = cluster ? P : sv.count[S];
        if (cluster) {
            if ((int)threadIdx.x < DIM * n1) {
                const int d = threadIdx.x / n1, b = threadIdx.x % n1;
                cn[d][b] = proxy_node(tabs.b[b], box_lo<DIM>(sv, d, S), box_hi<DIM>(sv, d, S));
            }
            __syncthreads();
        }
        for (int t0 = 0; t0 < nsrc; t0 += TB) {
            const int j = t0 + threadIdx.x;
            if (j < nsrc) tile[threadIdx.x] = load_src<DIM, F32>(sv, cluster, S, j, n1, P, cn, cx, cy, cz);
            __syncthreads();
            const int nt = min(TB, nsrc - t0);
            float a32[C];
#pragma unroll
            for (int c = 0; c < C; c++) a32[c] = 0.0f;
            for (int jj = g; jj < nt; jj += SPLIT)
                eval_any<DIM, KERNEL, F32>(tile[jj], t64, t32, kp.eps2, kp.excl, acc, a32);
            if constexpr (F32) {
#pragma unroll
                for (int c = 0; c < C; c++) acc[c] += (double)a32[c];
            }
            __syncthreads();
        }
    }
    if constexpr (SPLIT > 1) {
#pragma unroll
        for (int c = 0; c < C; c++) red[g][slot][c] = acc[c];
        __syncthreads();
        if (g == 0) {
            for (int gg = 1; gg < SPLIT; gg++)
#pragma unroll
                for (int c = 0; c < C; c++) acc[c] += red[gg][slot][c];
        }
    }
    if (g == 0 && idx < P) {
#pragma unroll
        for (int c = 0; c < C; c++)
            Psi[c * (size_t)pct * P + (size_t)T * P + idx] += scale_of<DIM, KERNEL, F32>(c, kp.gfc) * acc[c];
    }
}

// Work items: for every target panel with entries of the group, one item per
// block of `per` targets (n_targets = count of the leaf, or P for clusters).
__global__ void k_task_count(int pct, const int *__restrict__ goff, int far, const int *__restrict__ tcount, int P,
                             int per, int *__restrict__ ntask) {
    const int p = blockIdx.x * TB + threadIdx.x;
    if (p >= pct) return;
    const int ne = goff[far * pct + p + 1] - goff[far * pct + p];
    const int nt = far ? P : tcount[p];
    ntask[p] = ne > 0 ? (nt + per - 1) / per : 0;
}

__global__ void k_task_fill(int pct, const int *__restrict__ ntask, const int *__restrict__ tbase, int per,
                            int *__restrict__ task_panel, int *__restrict__ task_first) {
    const int p = blockIdx.x * TB + threadIdx.x;
    if (p >= pct) return;
    const int b = tbase[p];
    for (int j = 0; j < ntask[p]; j++) {
        task_panel[b + j] = p;
        task_first[b + j] = j * per;
    }
}

struct Tasks {
    int n = 0;
    DBuf<int> panel, first;
};

void build_tasks(const List &l, const Tree &tt, int far, int P, int per, Tasks &tk, hipStream_t s) {
    const int pct = tt.pc;
    DBuf<int> nt(pct), tb(pct);
    k_task_count<<<nblk(pct), TB, 0, s>>>(pct, l.goff.p, far, tt.count.p, P, per, nt.p);
    BTK_LAUNCH_CHECK();
    exclusive_sum_i32(nt.p, tb.p, pct, s);
    int *h = reinterpret_cast<int *>(host_scratch().v);
    BTK_HIP(hipMemcpyAsync(h, tb.p + (pct - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipMemcpyAsync(h + 1, nt.p + (pct - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    tk.n = h[0] + h[1];
    tk.panel.resize(tk.n ? tk.n : 1);
    tk.first.resize(tk.n ? tk.n : 1);
    k_task_fill<<<nblk(pct), TB, 0, s>>>(pct, nt.p, tb.p, per, tk.panel.p, tk.first.p);
    BTK_LAUNCH_CHECK();
}

template <int DIM, int KERNEL, bool F32, int TGT>
void launch_far(const Tasks &tk, const List &l, const Tree &tt, const SrcView &sv, int n1, const NodeTables &tabs,
                const KParams &kp, double *Psi, hipStream_t s) {
    k_far<DIM, KERNEL, F32, TGT><<<tk.n, TB, 0, s>>>(tk.panel.p, tk.first.p, l.goff.p, l.gsrc.p, l.gtype.p, tt.lo[0].p,
                                                     tt.lo[1].p, tt.lo[2].p, tt.hi[0].p, tt.hi[1].p, tt.hi[2].p, sv,
                                                     n1, tabs, kp, Psi, tt.pc);
}

template <int DIM, int KERNEL, bool F32>
void launch_interactions(const btk_run_config &cfg, const KParams &kp, const Tree &tt, const List &l,
                         const SrcView &sv, const NodeTables &tabs, double *out, double *Psi, hipStream_t s,
                         KernelEvents *kev) {
    const int n1 = cfg.interp_degree + 1;
    const int P = ipow(n1, DIM);
    if (l.counts[0] + l.counts[1] > 0) {
        constexpr int TPT = F32 ? 2 : 1;
        Tasks tk;
        build_tasks(l, tt, 0, P, NEAR_LANES * TPT, tk, s);
        if (kev) BTK_HIP(hipEventRecord(kev->ev[0], s));
        if (tk.n)
            k_near<DIM, KERNEL, F32, TPT><<<tk.n, NEAR_LANES, 0, s>>>(
                tk.panel.p, tk.first.p, l.goff.p, l.gsrc.p, l.gtype.p, tt.start.p, tt.count.p, tt.lo[0].p,
                tt.lo[1].p, tt.lo[2].p, tt.hi[0].p, tt.hi[1].p, tt.hi[2].p, tt.sx.p, tt.sy.p, tt.sz.p, tt.perm.p, sv,
                n1, tabs, kp, out, tt.n);
        BTK_LAUNCH_CHECK();
        if (kev) {
            BTK_HIP(hipEventRecord(kev->ev[1], s));
            kev->near_done = true;
        }
    }
    if (l.counts[2] + l.counts[3] > 0) {
        Tasks tk;
        const int tgt = P <= 64 ? 64 : (P <= 128 ? 128 : 256);
        build_tasks(l, tt, 1, P, tgt, tk, s);
        if (kev) BTK_HIP(hipEventRecord(kev->ev[2], s));
        if (tk.n) {
            if (tgt == 64)
                launch_far<DIM, KERNEL, F32, 64>(tk, l, tt, sv, n1, tabs, kp, Psi, s);
            else if (tgt == 128)
                launch_far<DIM, KERNEL, F32, 128>(tk, l, tt, sv, n1, tabs, kp, Psi, s);
            else
                launch_far<DIM, KERNEL, F32, 256>(tk, l, tt, sv, n1, tabs, kp, Psi, s);
        }
        BTK_LAUNCH_CHECK();
        if (kev) {
            BTK_HIP(hipEventRecord(kev->ev[3], s));
            kev->far_done = true;
        }
    }
}

template <int DIM, int KERNEL>
void launch_prec(const btk_run_config &cfg, const btk_options &o, const KParams &kp, const Tree &tt, const List &l,
                 const SrcView &sv, const NodeTables &tabs, double *out, double *Psi, hipStream_t s,
                 KernelEvents *kev) {
    if (o.precision == BTK_PREC_MIXED)
        launch_interactions<DIM, KERNEL, true>(cfg, kp, tt, l, sv, tabs, out, Psi, s, kev);
    else
        launch_interactions<DIM, KERNEL, false>(cfg, kp, tt, l, sv, tabs, out, Psi, s, kev);
}

// ---------------------------------------------------------------------------
// FP64 direct sum: one thread per target, sources streamed through LDS.
// ---------------------------------------------------------------------------
template <int DIM, int KERNEL>
__global__ void __launch_bounds__(TB)
    k_direct(int m, const int *__restrict__ idx, const double *__restrict__ xt, const double *__restrict__ yt,
             const double *__restrict__ zt, int n, const double *__restrict__ xs, const double *__restrict__ ys,
             const double *__restrict__ zs, const double *__restrict__ q, KParams kp, double *__restrict__ out) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    __shared__ double tile[4][TB];
    const int a = blockIdx.x * TB + threadIdx.x;
    const bool valid = a < m;
    const int i = valid ? (idx ? idx[a] : a) : 0;
    const double tx = valid ? xt[i] : 0.0, ty = valid ? yt[i] : 0.0;
    const double tz = (DIM == 3 && valid) ? zt[i] : 0.0;
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.0;
    for (int t0 = 0; t0 < n; t0 += TB) {
        const int j = t0 + threadIdx.x;
        if (j < n) {
            tile[0][threadIdx.x] = xs[j];
            tile[1][threadIdx.x] = ys[j];
            if (DIM == 3) tile[2][threadIdx.x] = zs[j];
            tile[3][threadIdx.x] = q[j];
        }
        __syncthreads();
        const int nt = min(TB, n - t0);
        for (int k = 0; k < nt; k++)
            eval_f64<DIM, KERNEL>(tx, ty, tz, tile[0][k], tile[1][k], DIM == 3 ? tile[2][k] : 0.0, tile[3][k], kp.eps2,
                                  kp.excl, acc);
        __syncthreads();
    }
    if (valid) finish_write<DIM, KERNEL>(acc, kp.gfc, out, (size_t)m, (size_t)a);
}

}  // namespace


void fmm_interactions(const btk_run_config &cfg, const btk_options &o, const Tree &tt, const Tree &ts, const List &l,
                      const double *q_sorted, const double *W, double *out, double *Psi, hipStream_t s,
                      KernelEvents *kev) {
    BTK_REQUIRE(tt.dim == ts.dim, "target and source trees differ in dimension");
    const int n1 = cfg.interp_degree + 1;
    BTK_REQUIRE(n1 >= 2 && n1 <= kMaxN1, "interp_degree must be between 1 and 10");
    BTK_REQUIRE(o.precision == BTK_PREC_FP64 || o.precision == BTK_PREC_MIXED, "unknown precision");
    const KParams kp = make_kparams(tt.dim, cfg.ker_eps, o);
    const NodeTables tabs = node_tables(cfg.interp_degree);
    const SrcView sv{ts.start.p, ts.count.p, ts.sx.p, ts.sy.p, ts.sz.p, q_sorted, ts.lo[0].p, ts.lo[1].p,
                     ts.lo[2].p, ts.hi[0].p, ts.hi[1].p, ts.hi[2].p, W};
    if (tt.dim == 3) {
        if (o.kernel == BTK_KERNEL_POT)
            launch_prec<3, BTK_KERNEL_POT>(cfg, o, kp, tt, l, sv, tabs, out, Psi, s, kev);
        else if (o.kernel == BTK_KERNEL_REF_VEL3D)
            launch_prec<3, BTK_KERNEL_REF_VEL3D>(cfg, o, kp, tt, l, sv, tabs, out, Psi, s, kev);
        else
            launch_prec<3, BTK_KERNEL_POT_GRAD>(cfg, o, kp, tt, l, sv, tabs, out, Psi, s, kev);
    } else {
        BTK_REQUIRE(o.kernel != BTK_KERNEL_REF_VEL3D, "the reference vel kernel is 3D only");
        if (o.kernel == BTK_KERNEL_POT)
            launch_prec<2, BTK_KERNEL_POT>(cfg, o, kp, tt, l, sv, tabs, out, Psi, s, kev);
        else
            launch_prec<2, BTK_KERNEL_POT_GRAD>(cfg, o, kp, tt, l, sv, tabs, out, Psi, s, kev);
    }
}

void direct_sum(int dim, const btk_options &o, double eps, int m, const int *idx, const double *xt, const double *yt,
                const double *zt, int n, const double *xs, const double *ys, const double *zs, const double *q,
                double *out, hipStream_t s) {
    if (m <= 0) return;
    const KParams kp = make_kparams(dim, eps, o);
    const int g = nblk(m);
    if (dim == 3) {
        if (o.kernel == BTK_KERNEL_POT)
            k_direct<3, BTK_KERNEL_POT><<<g, TB, 0, s>>>(m, idx, xt, yt, zt, n, xs, ys, zs, q, kp, out);
        else if (o.kernel == BTK_KERNEL_REF_VEL3D)
            k_direct<3, BTK_KERNEL_REF_VEL3D><<<g, TB, 0, s>>>(m, idx, xt, yt, zt, n, xs, ys, zs, q, kp, out);
        else
            k_direct<3, BTK_KERNEL_POT_GRAD><<<g, TB, 0, s>>>(m, idx, xt, yt, zt, n, xs, ys, zs, q, kp, out);
    } else {
        BTK_REQUIRE(o.kernel != BTK_KERNEL_REF_VEL3D, "the reference vel kernel is 3D only");
        if (o.kernel == BTK_KERNEL_POT)
            k_direct<2, BTK_KERNEL_POT><<<g, TB, 0, s>>>(m, idx, xt, yt, nullptr, n, xs, ys, nullptr, q, kp, out);
        else
            k_direct<2, BTK_KERNEL_POT_GRAD><<<g, TB, 0, s>>>(m, idx, xt, yt, nullptr, n, xs, ys, nullptr, q, kp,
                                                               out);
    }
    BTK_LAUNCH_CHECK();
}

}  // namespace btk
