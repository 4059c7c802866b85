// PP / PC / CP / CC interactions (poisson_fmm.cpp:8-26, poisson_fmm_impl.hpp:
// 10-442; poisson_vel_fmm.cpp:8-17, poisson_vel_fmm_impl.hpp:10-257) and the
// FP64 direct sum (direct_sum_impl.hpp:9-94).
//
// The reference runs one thread per interaction pair and adds every kernel
// evaluation to its target with an atomic.  Here the list is grouped by
// target (stably, so each target sees its sources in the reference's list
// order) and every target is owned by one thread:
//   near kernel  a workgroup owns a block of up to TBT points of one target
//                leaf and streams the sources of the leaf's PP pairs (the
//                source leaf's points) and PC pairs (the source cluster's
//                proxy points and weights) through LDS tiles;
//   far kernel   a workgroup owns a block of the proxy points of one target
//                cluster and streams the sources of its CP pairs (points) and
//                CC pairs (proxy points).
// Source tiles are coalesced SoA loads of the tree-sorted coordinates; each
// thread reads every tile entry as an LDS broadcast.  Results are written
// once per target, with the kernel's constant factor applied.
#include "btk_device.hpp"

namespace btk {
namespace {

constexpr int TB = 256;
inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }

struct SrcView {
    const int *start, *count;
    const double *sx, *sy, *sz, *q;  // sorted points and charges
    const double *lo0, *lo1, *lo2, *hi0, *hi1, *hi2;
    const double *W;  // [pc][P]
};

template <int DIM, int KERNEL>
__device__ __forceinline__ void finish_write(const double *acc, double gfc, double *out, size_t stride, size_t at) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
#pragma unroll
    for (int c = 0; c < C; c++) out[c * stride + at] += comp_scale<DIM, KERNEL>(c, gfc) * acc[c];
}

// Stream the sources of one list entry through LDS and accumulate into acc.
template <int DIM, int KERNEL, int TBT>
__device__ __forceinline__ void stream_entry(int S, int type_is_cluster, const SrcView &sv, int n1,
                                             const NodeTables &tabs, double tx, double ty, double tz, double eps2,
                                             int excl, double *acc, double (*tile)[TBT], double (*cn)[kMaxN1]) {
    if (!type_is_cluster) {
        const int ns = sv.count[S], ss = sv.start[S];
        for (int t0 = 0; t0 < ns; t0 += TBT) {
            const int j = t0 + threadIdx.x;
            if (j < ns) {
                tile[0][threadIdx.x] = sv.sx[ss + j];
                tile[1][threadIdx.x] = sv.sy[ss + j];
                if (DIM == 3) tile[2][threadIdx.x] = sv.sz[ss + j];
                tile[3][threadIdx.x] = sv.q[ss + j];
            }
            __syncthreads();
            const int nt = min(TBT, ns - t0);
            for (int k = 0; k < nt; k++)
                eval_f64<DIM, KERNEL>(tx, ty, tz, tile[0][k], tile[1][k], DIM == 3 ? tile[2][k] : 0.0, tile[3][k],
                                      eps2, excl, acc);
            __syncthreads();
        }
    } else {
        const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
        if ((int)threadIdx.x < DIM * n1) {
            const int d = threadIdx.x / n1, b = threadIdx.x % n1;
            const double lo = d == 0 ? sv.lo0[S] : (d == 1 ? sv.lo1[S] : sv.lo2[S]);
            const double hi = d == 0 ? sv.hi0[S] : (d == 1 ? sv.hi1[S] : sv.hi2[S]);
            cn[d][b] = proxy_node(tabs.b[b], lo, hi);  // bli_points_shift (poisson_fmm_impl.hpp:292-294)
        }
        __syncthreads();
        const double *w = sv.W + (size_t)S * P;
        for (int t0 = 0; t0 < P; t0 += TBT) {
            const int idx = t0 + threadIdx.x;
            if (idx < P) {
                if (DIM == 3) {
                    const int l = idx % n1, k = (idx / n1) % n1, j = idx / (n1 * n1);
                    tile[0][threadIdx.x] = cn[0][j];
                    tile[1][threadIdx.x] = cn[1][k];
                    tile[2][threadIdx.x] = cn[2][l];
                } else {
                    const int k = idx % n1, j = idx / n1;
                    tile[0][threadIdx.x] = cn[0][j];
                    tile[1][threadIdx.x] = cn[1][k];
                }
                tile[3][threadIdx.x] = w[idx];
            }
            __syncthreads();
            const int nt = min(TBT, P - t0);
            for (int k = 0; k < nt; k++)
                eval_f64<DIM, KERNEL>(tx, ty, tz, tile[0][k], tile[1][k], DIM == 3 ? tile[2][k] : 0.0, tile[3][k],
                                      eps2, excl, acc);
            __syncthreads();
        }
    }
}

// PP + PC for a block of points of one target leaf.
template <int DIM, int KERNEL, int TBT>
__global__ void __launch_bounds__(TBT)
    k_near(const int *__restrict__ task_panel, const int *__restrict__ task_first, const int *__restrict__ goff,
           const int *__restrict__ gsrc, const unsigned char *__restrict__ gtype, const int *__restrict__ tstart,
           const int *__restrict__ tcount, const double *__restrict__ tx_, const double *__restrict__ ty_,
           const double *__restrict__ tz_, const int *__restrict__ tperm, SrcView sv, int n1, NodeTables tabs,
           KParams kp, double *__restrict__ out, int M) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    __shared__ double tile[4][TBT];
    __shared__ double cn[3][kMaxN1];
    const int T = task_panel[blockIdx.x];
    const int first = task_first[blockIdx.x];
    const int ti = first + threadIdx.x;
    const bool valid = ti < tcount[T];
    const int i = tstart[T] + (valid ? ti : 0);
    const double tx = tx_[i], ty = ty_[i], tz = DIM == 3 ? tz_[i] : 0.0;
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.0;
    const int e0 = goff[T], e1 = goff[T + 1];
    for (int e = e0; e < e1; e++)
        stream_entry<DIM, KERNEL, TBT>(gsrc[e], gtype[e] == 1, sv, n1, tabs, tx, ty, tz, kp.eps2, kp.excl, acc, tile,
                                       cn);
    if (valid) finish_write<DIM, KERNEL>(acc, kp.gfc, out, (size_t)M, (size_t)tperm[i]);
}

// CP + CC for a block of proxy points of one target cluster.
template <int DIM, int KERNEL, int TBT>
__global__ void __launch_bounds__(TBT)
    k_far(const int *__restrict__ task_panel, const int *__restrict__ task_first, const int *__restrict__ goff,
          const int *__restrict__ gsrc, const unsigned char *__restrict__ gtype, const double *__restrict__ tlo0,
          const double *__restrict__ tlo1, const double *__restrict__ tlo2, const double *__restrict__ thi0,
          const double *__restrict__ thi1, const double *__restrict__ thi2, SrcView sv, int n1, NodeTables tabs,
          KParams kp, double *__restrict__ Psi, int pct) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    __shared__ double tile[4][TBT];
    __shared__ double cn[3][kMaxN1];
    const int T = task_panel[blockIdx.x];
    const int first = task_first[blockIdx.x];
    const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
    const int idx = first + threadIdx.x;
    const bool valid = idx < P;
    const int id = valid ? idx : 0;
    // target proxy point (poisson_fmm_impl.hpp:352-361)
    double tx, ty, tz = 0.0;
    if (DIM == 3) {
        const int l = id % n1, k = (id / n1) % n1, j = id / (n1 * n1);
        tx = proxy_node(tabs.b[j], tlo0[T], thi0[T]);
        ty = proxy_node(tabs.b[k], tlo1[T], thi1[T]);
        tz = proxy_node(tabs.b[l], tlo2[T], thi2[T]);
    } else {
        const int k = id % n1, j = id / n1;
        tx = proxy_node(tabs.b[j], tlo0[T], thi0[T]);
        ty = proxy_node(tabs.b[k], tlo1[T], thi1[T]);
    }
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.0;
    const int e0 = goff[pct + T], e1 = goff[pct + T + 1];
    for (int e = e0; e < e1; e++)
        stream_entry<DIM, KERNEL, TBT>(gsrc[e], gtype[e] == 3, sv, n1, tabs, tx, ty, tz, kp.eps2, kp.excl, acc, tile,
                                       cn);
    if (valid) finish_write<DIM, KERNEL>(acc, kp.gfc, Psi, (size_t)pct * P, (size_t)T * P + idx);
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

template <int DIM, int KERNEL>
void launch_interactions(const btk_run_config &cfg, const KParams &kp, const Tree &tt, const Tree &ts, const List &l,
                         const SrcView &sv, const NodeTables &tabs, double *out, double *Psi, hipStream_t s,
                         KernelEvents *kev) {
    const int n1 = cfg.interp_degree + 1;
    const int P = ipow(n1, DIM);
    if (l.counts[0] + l.counts[1] > 0) {
        Tasks tk;
        build_tasks(l, tt, 0, P, 256, tk, s);
        if (kev) BTK_HIP(hipEventRecord(kev->ev[0], s));
        if (tk.n)
            k_near<DIM, KERNEL, 256><<<tk.n, 256, 0, s>>>(tk.panel.p, tk.first.p, l.goff.p, l.gsrc.p, l.gtype.p,
                                                         tt.start.p, tt.count.p, tt.sx.p, tt.sy.p, tt.sz.p, tt.perm.p,
                                                         sv, n1, tabs, kp, out, tt.n);
        BTK_LAUNCH_CHECK();
        if (kev) {
            BTK_HIP(hipEventRecord(kev->ev[1], s));
            kev->near_done = true;
        }
    }
    if (l.counts[2] + l.counts[3] > 0) {
        Tasks tk;
        if (P <= 64) {
            build_tasks(l, tt, 1, P, 64, tk, s);
            if (kev) BTK_HIP(hipEventRecord(kev->ev[2], s));
            if (tk.n)
                k_far<DIM, KERNEL, 64><<<tk.n, 64, 0, s>>>(tk.panel.p, tk.first.p, l.goff.p, l.gsrc.p, l.gtype.p,
                                                          tt.lo[0].p, tt.lo[1].p, tt.lo[2].p, tt.hi[0].p, tt.hi[1].p,
                                                          tt.hi[2].p, sv, n1, tabs, kp, Psi, tt.pc);
        } else {
            build_tasks(l, tt, 1, P, 256, tk, s);
            if (kev) BTK_HIP(hipEventRecord(kev->ev[2], s));
            if (tk.n)
                k_far<DIM, KERNEL, 256><<<tk.n, 256, 0, s>>>(tk.panel.p, tk.first.p, l.goff.p, l.gsrc.p, l.gtype.p,
                                                            tt.lo[0].p, tt.lo[1].p, tt.lo[2].p, tt.hi[0].p,
                                                            tt.hi[1].p, tt.hi[2].p, sv, n1, tabs, kp, Psi, tt.pc);
        }
        BTK_LAUNCH_CHECK();
        if (kev) {
            BTK_HIP(hipEventRecord(kev->ev[3], s));
            kev->far_done = true;
        }
    }
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
    const KParams kp = make_kparams(tt.dim, cfg.ker_eps, o);
    const NodeTables tabs = node_tables(cfg.interp_degree);
    const SrcView sv{ts.start.p, ts.count.p, ts.sx.p, ts.sy.p, ts.sz.p, q_sorted, ts.lo[0].p, ts.lo[1].p,
                     ts.lo[2].p, ts.hi[0].p, ts.hi[1].p, ts.hi[2].p, W};
    if (tt.dim == 3) {
        if (o.kernel == BTK_KERNEL_POT)
            launch_interactions<3, BTK_KERNEL_POT>(cfg, kp, tt, ts, l, sv, tabs, out, Psi, s, kev);
        else if (o.kernel == BTK_KERNEL_REF_VEL3D)
            launch_interactions<3, BTK_KERNEL_REF_VEL3D>(cfg, kp, tt, ts, l, sv, tabs, out, Psi, s, kev);
        else
            launch_interactions<3, BTK_KERNEL_POT_GRAD>(cfg, kp, tt, ts, l, sv, tabs, out, Psi, s, kev);
    } else {
        BTK_REQUIRE(o.kernel != BTK_KERNEL_REF_VEL3D, "the reference vel kernel is 3D only");
        if (o.kernel == BTK_KERNEL_POT)
            launch_interactions<2, BTK_KERNEL_POT>(cfg, kp, tt, ts, l, sv, tabs, out, Psi, s, kev);
        else
            launch_interactions<2, BTK_KERNEL_POT_GRAD>(cfg, kp, tt, ts, l, sv, tabs, out, Psi, s, kev);
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
