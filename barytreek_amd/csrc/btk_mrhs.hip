// Many charge vectors over one geometry (no reference counterpart; the
// reference solves one charge vector per run).
//
// Tree, list, basis matrices and kernel evaluations are shared by all K
// vectors, so the interactions become products G Q of the kernel tile G
// (targets x sources) with the charge tile Q (sources x K):
//   a workgroup of 4 wavefronts owns 64 targets of one panel and 64 charge
//   vectors; wave w takes targets 32*(w&1) .. +32 and vectors 32*(w>>1) .. +32;
//   per 64-source LDS tile every lane evaluates G for its target and one
//   source parity (FP64 difference rounded to FP32, kernel in FP32), which is
//   exactly the A operand of v_mfma_f32_32x32x2f32, while the B operand is
//   read from the charge tile in LDS;
//   the FP32 accumulators of each tile are added into FP64 per-target sums.
// Operand layouts of v_mfma_f32_32x32x2f32 (checked on the device): lane l
// holds A[l%32][l/32], B[l/32][l%32], and D[8*(v/4) + 4*(l/32) + v%4][l%32]
// in accumulator register v.
#include "btk_interact.hpp"

namespace btk {
namespace {

constexpr int TB = 256;
constexpr int KC = 64;  // charge vectors per workgroup
inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }

typedef float f32x16 __attribute__((ext_vector_type(16)));

// G components (without the constant factor) of one pair; R as in eval_f32.
template <int DIM, int KERNEL, bool EXCL>
__device__ __forceinline__ void kernel_row(float dx, float dy, float dz, float eps2, float *g) {
    if constexpr (DIM == 3) {
        const float R = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const bool skip = EXCL && R == eps2;
        const float inv = skip ? 0.0f : __builtin_amdgcn_rsqf(R);
        g[0] = inv;
        if constexpr (KERNEL == BTK_KERNEL_POT_GRAD) {
            const float h = inv * inv * inv;
            g[1] = h * dx;
            g[2] = h * dy;
            g[3] = h * dz;
        }
    } else {
        (void)dz;
        const float R = fmaf(dx, dx, fmaf(dy, dy, eps2));
        const bool skip = EXCL && R == eps2;
        g[0] = skip ? 0.0f : __builtin_amdgcn_logf(R);
        if constexpr (KERNEL == BTK_KERNEL_POT_GRAD) {
            const float r = skip ? 0.0f : __builtin_amdgcn_rcpf(R);
            g[1] = r * dx;
            g[2] = r * dy;
        }
    }
}

struct MView {
    const float *q;    // [n_sources][kp]  sorted source charges, FP32
    const float *w;    // [pc_s * P][kp]   proxy source weights, FP32
    double *out;       // near: [K][C][M]; far: [K][C][pc_t][P]
    int K, kp;
};

template <int DIM, int KERNEL, bool FAR, bool EXCL>
__global__ void __launch_bounds__(TB)
    k_interact_mrhs(TaskView tk, GroupView gv, TgtView tv, SrcView sv, MView mv, int n1, NodeTables tabs,
                    KParams kp, int M, int pct, int atomic, int same_tree) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    __shared__ double sc[3][64];
    __shared__ float qt[64][KC + 1];
    __shared__ double cn[3][kMaxN1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int item = xcd_contiguous(blockIdx.x, gridDim.x);
    const int T = tk.panel[item];
    const int first = tk.first[item];
    const int e0 = tk.e0[item], e1 = tk.e1[item];
    const int rhs0 = blockIdx.y * KC;
    const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
    const int ntgt = FAR ? P : tv.count[T];
    const int ts0 = FAR ? 0 : tv.start[T];
    // this lane's target (A rows) and charge-vector column (B, D columns)
    const int ti = first + 32 * (w & 1) + (lane & 31);
    const int tc = min(ti, ntgt - 1);
    double t[3];
    if constexpr (FAR) {
        proxy_point<DIM>(tc, n1, tabs, tv.lo0[T], tv.hi0[T], tv.lo1[T], tv.hi1[T], DIM == 3 ? tv.lo2[T] : 0.0,
                         DIM == 3 ? tv.hi2[T] : 0.0, t);
    } else {
        t[0] = tv.sx[ts0 + tc];
        t[1] = tv.sy[ts0 + tc];
        t[2] = DIM == 3 ? tv.sz[ts0 + tc] : 0.0;
    }
    const int col = 32 * (w >> 1) + (lane & 31);
    const int par = lane >> 5;
    double acc[C][16];
#pragma unroll
    for (int c = 0; c < C; c++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[c][v] = 0.0;
    const float eps2 = (float)kp.eps2;
    for (int e = e0; e < e1; e++) {
        const int S = gv.src[e];
        const bool cluster = (gv.type[e] & 1) != 0;
        const int nsrc = cluster ? P : sv.count[S];
        if (cluster) {
            if (tid < DIM * n1) {
                const int d = tid / n1, b = tid % n1;
                cn[d][b] = proxy_node(tabs.b[b], sel3(d, sv.lo0[S], sv.lo1[S], sv.lo2[S]),
                                      sel3(d, sv.hi0[S], sv.hi1[S], sv.hi2[S]));
            }
            __syncthreads();
        }
        const bool check = EXCL && (!same_tree || FAR || cluster || S == T);
        const int ss = cluster ? 0 : sv.start[S];
        for (int t0 = 0; t0 < nsrc; t0 += 64) {
            const int nt = min(64, nsrc - t0);
            if (tid < 64) {
                const int j = t0 + tid;
                double x = 0.0, y = 0.0, z = 0.0;  // padding sources: their G is zeroed below
                if (tid < nt) {
                    if (cluster) {
                        if constexpr (DIM == 3) {
                            x = cn[0][j / (n1 * n1)];
                            y = cn[1][(j / n1) % n1];
                            z = cn[2][j % n1];
                        } else {
                            x = cn[0][j / n1];
                            y = cn[1][j % n1];
                            z = 0.0;
                        }
                    } else {
                        x = sv.sx[ss + j];
                        y = sv.sy[ss + j];
                        z = DIM == 3 ? sv.sz[ss + j] : 0.0;
                    }
                }
                sc[0][tid] = x;
                sc[1][tid] = y;
                sc[2][tid] = z;
            }
            const float *qsrc = cluster ? mv.w + ((size_t)S * P + t0) * mv.kp : mv.q + (size_t)(ss + t0) * mv.kp;
            for (int k = tid; k < 64 * KC; k += TB) {
                const int s = k / KC, r = k % KC;
                qt[s][r] = (s < nt) ? qsrc[(size_t)s * mv.kp + rhs0 + r] : 0.0f;
            }
            __syncthreads();
            f32x16 d[C];
#pragma unroll
            for (int c = 0; c < C; c++) d[c] = (f32x16){0};
#pragma unroll 4
            for (int step = 0; step < 32; step++) {
                const int s = 2 * step + par;
                const float dx = (float)(t[0] - sc[0][s]);
                const float dy = (float)(t[1] - sc[1][s]);
                const float dz = DIM == 3 ? (float)(t[2] - sc[2][s]) : 0.0f;
                float g[C];
                if (check)
                    kernel_row<DIM, KERNEL, true>(dx, dy, dz, eps2, g);
                else
                    kernel_row<DIM, KERNEL, false>(dx, dy, dz, eps2, g);
                if (s >= nt) {
#pragma unroll
                    for (int c = 0; c < C; c++) g[c] = 0.0f;
                }
                const float b = qt[s][col];
#pragma unroll
                for (int c = 0; c < C; c++) d[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(g[c], b, d[c], 0, 0, 0);
            }
#pragma unroll
            for (int c = 0; c < C; c++)
#pragma unroll
                for (int v = 0; v < 16; v++) acc[c][v] += (double)d[c][v];
            __syncthreads();
        }
    }
    // D[i][j]: target row i = 8*(v/4) + 4*(lane/32) + v%4, vector column j = lane%32
    const int r = rhs0 + col;
    if (r >= mv.K) return;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int row = 8 * (v / 4) + 4 * par + (v % 4);
        const int tt = first + 32 * (w & 1) + row;
        if (tt >= ntgt) continue;
#pragma unroll
        for (int c = 0; c < C; c++) {
            double scale = comp_scale_f32<DIM, KERNEL>(c, kp.gfc);
            const double val = scale * acc[c][v];
            double *dst;
            if constexpr (FAR)
                dst = mv.out + (((size_t)r * C + c) * pct + T) * P + tt;
            else
                dst = mv.out + ((size_t)r * C + c) * M + tv.perm[ts0 + tt];
            if (atomic)
                unsafeAtomicAdd(dst, val);
            else
                *dst += val;
        }
    }
}

__global__ void k_gather_multi(int n, int K, int kp, const int *__restrict__ perm, const double *__restrict__ q,
                               float *__restrict__ qs) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    for (int r = 0; r < kp; r++) qs[(size_t)i * kp + r] = r < K ? (float)q[(size_t)r * n + o] : 0.0f;
}

// W[K][rows] (FP64) -> Wt[rows][kp] (FP32)
__global__ void k_transpose_w(long long rows, int K, int kp, const double *__restrict__ W, float *__restrict__ Wt) {
    const long long i = (long long)blockIdx.x * TB + threadIdx.x;
    if (i >= rows) return;
    for (int r = 0; r < kp; r++) Wt[i * kp + r] = r < K ? (float)W[(size_t)r * rows + i] : 0.0f;
}

}  // namespace

// Declared in btk_internal.hpp.
void gather_multi_f32(const Tree &t, const double *q, int K, int kp, float *qs, hipStream_t s) {
    if (t.n == 0) return;
    k_gather_multi<<<nblk(t.n), TB, 0, s>>>(t.n, K, kp, t.perm.p, q, qs);
    BTK_LAUNCH_CHECK();
}

void transpose_weights_f32(long long rows, int K, int kp, const double *W, float *Wt, hipStream_t s) {
    if (rows == 0) return;
    k_transpose_w<<<nblk(rows), TB, 0, s>>>(rows, K, kp, W, Wt);
    BTK_LAUNCH_CHECK();
}

template <int DIM, int KERNEL>
void launch_mrhs(const TaskView &tk, int n_tasks, const GroupView &gv, const TgtView &tv, const SrcView &sv,
                 const float *qs, const float *wt, double *out, int K, int kp, int n1, const NodeTables &tabs,
                 const KParams &kp_, int M, int pct, int atomic, int same_tree, bool far, bool excl, hipStream_t s) {
    if (n_tasks <= 0) return;
    const MView mv{qs, wt, out, K, kp};
    const dim3 grid(n_tasks, kp / KC);
    if (far) {
        if (excl)
            k_interact_mrhs<DIM, KERNEL, true, true><<<grid, TB, 0, s>>>(tk, gv, tv, sv, mv, n1, tabs, kp_, M, pct,
                                                                        atomic, same_tree);
        else
            k_interact_mrhs<DIM, KERNEL, true, false><<<grid, TB, 0, s>>>(tk, gv, tv, sv, mv, n1, tabs, kp_, M, pct,
                                                                         atomic, same_tree);
    } else {
        if (excl)
            k_interact_mrhs<DIM, KERNEL, false, true><<<grid, TB, 0, s>>>(tk, gv, tv, sv, mv, n1, tabs, kp_, M, pct,
                                                                         atomic, same_tree);
        else
            k_interact_mrhs<DIM, KERNEL, false, false><<<grid, TB, 0, s>>>(tk, gv, tv, sv, mv, n1, tabs, kp_, M,
                                                                          pct, atomic, same_tree);
    }
    BTK_LAUNCH_CHECK();
}

template <int DIM, int KERNEL>
void run_multi(const btk_run_config &cfg, const btk_options &o, const Tree &tt, const Tree &ts, const List &l,
               const float *qs, const float *wt, int K, int kp, double *out, double *Psi, hipStream_t s,
               KernelEvents *kev) {
    const int n1 = cfg.interp_degree + 1;
    const int P = ipow(n1, DIM);
    const KParams kp_ = make_kparams(DIM, cfg.ker_eps, o);
    const NodeTables tabs = node_tables(cfg.interp_degree);
    const SrcView sv{ts.start.p, ts.count.p, ts.sx.p, ts.sy.p, ts.sz.p, nullptr, ts.lo[0].p, ts.lo[1].p,
                     ts.lo[2].p, ts.hi[0].p, ts.hi[1].p, ts.hi[2].p, nullptr};
    const TgtView tv{tt.start.p, tt.count.p, tt.perm.p, tt.sx.p, tt.sy.p, tt.sz.p,
                     tt.lo[0].p, tt.lo[1].p, tt.lo[2].p, tt.hi[0].p, tt.hi[1].p, tt.hi[2].p};
    const GroupView gv{l.gsrc.p, l.gtype.p};
    const int chunks = kp / KC;
    for (int far = 0; far < 2; far++) {
        if (l.counts[2 * far] + l.counts[2 * far + 1] == 0) continue;
        const int n_clusters = tt.pc - tt.n_leaves;
        // base counts wavefronts (workgroups of 4): aim at 8192 of them
        const long long base = (far ? (long long)n_clusters * ((P + 63) / 64) : (long long)tt.n / 64 + tt.n_leaves) *
                               chunks * 4;
        const int nsplit = choose_split(base, 8192);
        Tasks tk;
        build_tasks(l, tt, ts, far, P, 64, nsplit, tk, s);
        if (kev) BTK_HIP(hipEventRecord(kev->ev[far ? 2 : 0], s));
        launch_mrhs<DIM, KERNEL>(TaskView{tk.panel.p, tk.first.p, tk.e0.p, tk.e1.p}, tk.n, gv, tv, sv, qs, wt,
                                 far ? Psi : out, K, kp, n1, tabs, kp_, tt.n, tt.pc, nsplit > 1 ? 1 : 0,
                                 (&tt == &ts) ? 1 : 0, far != 0, o.exclude_self != 0, s);
        if (kev) {
            BTK_HIP(hipEventRecord(kev->ev[far ? 3 : 1], s));
            (far ? kev->far_done : kev->near_done) = true;
        }
    }
}

// Declared in btk_internal.hpp.
void fmm_interactions_multi(const btk_run_config &cfg, const btk_options &o, const Tree &tt, const Tree &ts,
                            const List &l, const float *qs, const float *wt, int K, int kp, double *out, double *Psi,
                            hipStream_t s, KernelEvents *kev) {
    BTK_REQUIRE(kp % KC == 0 && kp >= K, "charge-vector padding must be a multiple of 64");
    BTK_REQUIRE(o.kernel != BTK_KERNEL_REF_VEL3D, "many charge vectors: potential or potential+gradient kernels");
    BTK_REQUIRE(l.ndeg == 1, "many charge vectors: one interpolation degree per list");
    if (tt.dim == 3) {
        if (o.kernel == BTK_KERNEL_POT)
            run_multi<3, BTK_KERNEL_POT>(cfg, o, tt, ts, l, qs, wt, K, kp, out, Psi, s, kev);
        else
            run_multi<3, BTK_KERNEL_POT_GRAD>(cfg, o, tt, ts, l, qs, wt, K, kp, out, Psi, s, kev);
    } else {
        if (o.kernel == BTK_KERNEL_POT)
            run_multi<2, BTK_KERNEL_POT>(cfg, o, tt, ts, l, qs, wt, K, kp, out, Psi, s, kev);
        else
            run_multi<2, BTK_KERNEL_POT_GRAD>(cfg, o, tt, ts, l, qs, wt, K, kp, out, Psi, s, kev);
    }
}

template void launch_mrhs<2, BTK_KERNEL_POT>(const TaskView &, int, const GroupView &, const TgtView &,
                                              const SrcView &, const float *, const float *, double *, int, int, int,
                                              const NodeTables &, const KParams &, int, int, int, int, bool, bool,
                                              hipStream_t);
template void launch_mrhs<2, BTK_KERNEL_POT_GRAD>(const TaskView &, int, const GroupView &, const TgtView &,
                                                   const SrcView &, const float *, const float *, double *, int, int,
                                                   int, const NodeTables &, const KParams &, int, int, int, int, bool,
                                                   bool, hipStream_t);
template void launch_mrhs<3, BTK_KERNEL_POT>(const TaskView &, int, const GroupView &, const TgtView &,
                                              const SrcView &, const float *, const float *, double *, int, int, int,
                                              const NodeTables &, const KParams &, int, int, int, int, bool, bool,
                                              hipStream_t);
template void launch_mrhs<3, BTK_KERNEL_POT_GRAD>(const TaskView &, int, const GroupView &, const TgtView &,
                                                   const SrcView &, const float *, const float *, double *, int, int,
                                                   int, const NodeTables &, const KParams &, int, int, int, int, bool,
                                                   bool, hipStream_t);

}  // namespace btk
