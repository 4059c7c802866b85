// Mutual PP interactions of one tree (mixed precision, opt-in launch shape).
//
// With one tree for targets and sources, the list holds most leaf pairs in
// both orders: (T, S) and (S, T) are both PP.  The kernel value of a point
// pair is symmetric (the gradient antisymmetric), so one evaluation of
// G(t - s) serves t (from s) and s (from t).  Such pairs are evaluated here
// once, for both directions; their list entries are then skipped by the
// target-grouped kernel (their source count in the entry descriptor is set
// to zero), which keeps the diagonal (T, T) and any one-way entries.
//
// Each unordered pair is owned by one of its two entries (alternating by
// panel parity, so that every leaf owns about half of its pairs).  A work
// item is one workgroup of W wavefronts: up to W*64*TPT targets of a leaf T
// against the owned partners S of a slice of T's entries.  Tiles of 64
// points of S sit in LDS; at step k lane i meets point (i + k) mod 64, and
// the per-point partial sums move one lane down after every step (DPP
// full-wave rotate), so that after 64 steps lane i holds the sums of point
// i.  The W wavefronts' sums of a point meet in LDS and go to the FP64
// output with one atomic add per component; the targets' sums stay in
// registers over all partners.
//
// Half the transcendentals and ~30% fewer VALU instructions per evaluation
// than the one-way kernel, but the workgroup barriers of the cross-wave
// reduction cost about as much on large leaves; the autotuner decides.
#include <algorithm>
#include <cmath>

#include "btk_interact.hpp"

namespace btk {
namespace {

constexpr int TB = 256;
inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }

// Panel of near entry e (goff is ascending).
__device__ __forceinline__ int entry_panel(const int *__restrict__ goff, int pct, int e) {
    int lo = 0, hi = pct;  // goff[lo] <= e < goff[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (goff[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The pair {T, S} belongs to T's entry when (T + S) is odd and T < S, or
// (T + S) is even and T > S.
__device__ __forceinline__ bool owns(int T, int S) { return ((T + S) & 1) ? T < S : T > S; }

// For every near entry e = (T, S): if it is PP with S != T and S's near
// entries hold PP (S, T), its source count in meta becomes 0 (the mutual
// kernel evaluates it) and, for the owning entry, mmeta[e] describes S's
// points; every other mmeta entry has no sources.
__global__ void k_mutual_mark(int nn, int pct, const int *__restrict__ goff, const int *__restrict__ gsrc,
                              const unsigned char *__restrict__ gtype, const int *__restrict__ start,
                              const int *__restrict__ count, int4 *__restrict__ meta, int4 *__restrict__ mmeta) {
    const int e = blockIdx.x * TB + threadIdx.x;
    if (e >= nn) return;
    const int S = gsrc[e];
    int4 mm = make_int4(S, 0, 0, 0);
    if (gtype[e] == 0) {
        const int T = entry_panel(goff, pct, e);
        if (S != T) {
            bool mutual = false;
            for (int g = goff[S]; g < goff[S + 1]; g++)
                if (gtype[g] == 0 && gsrc[g] == T) {
                    mutual = true;
                    break;
                }
            if (mutual) {
                meta[e].z = 0;
                if (owns(T, S)) mm = make_int4(S, start[S], count[S], 0);
            }
        }
    }
    mmeta[e] = mm;
}

// Lane i takes lane (i + 1) mod 64's value (DPP full-wave rotate).
__device__ __forceinline__ float rol1(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xf, 0xf, false));
}

// One point pair, both directions: at += q_s G(t - s) (and its gradient at
// t), as += q_t G(s - t) (and its gradient at s).  The same FP32 arithmetic
// as eval_f32, with d = t - s formed in FP64 and rounded once.
template <int DIM, int KERNEL>
__device__ __forceinline__ void eval_both(const double *t, float qt, double sx, double sy, double sz, float qs,
                                          float eps2, float *at, float *as) {
    const float dx = (float)(t[0] - sx), dy = (float)(t[1] - sy), dz = DIM == 3 ? (float)(t[2] - sz) : 0.0f;
    if constexpr (DIM == 3) {
        const float R = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        if constexpr (KERNEL == BTK_KERNEL_POT) {
            const float inv = __builtin_amdgcn_rsqf(R);
            at[0] = fmaf(qs, inv, at[0]);
            as[0] = fmaf(qt, inv, as[0]);
        } else if constexpr (KERNEL == BTK_KERNEL_REF_VEL3D) {
            const float rc = __builtin_amdgcn_rcpf(R);
            const float g = qs * rc, h = qt * rc;
            at[0] = fmaf(dx, g, at[0]);
            at[1] = fmaf(dy, g, at[1]);
            at[2] = fmaf(dz, g, at[2]);
            as[0] = fmaf(-dx, h, as[0]);
            as[1] = fmaf(-dy, h, as[1]);
            as[2] = fmaf(-dz, h, as[2]);
        } else {
            const float inv = __builtin_amdgcn_rsqf(R);
            const float i3 = inv * inv;
            const float qi = qs * inv, qj = qt * inv;
            at[0] += qi;
            as[0] += qj;
            const float g = qi * i3, h = qj * i3;
            at[1] = fmaf(g, dx, at[1]);
            at[2] = fmaf(g, dy, at[2]);
            at[3] = fmaf(g, dz, at[3]);
            as[1] = fmaf(-h, dx, as[1]);
            as[2] = fmaf(-h, dy, as[2]);
            as[3] = fmaf(-h, dz, as[3]);
        }
    } else {
        const float R = fmaf(dx, dx, fmaf(dy, dy, eps2));
        const float lg = __builtin_amdgcn_logf(R);
        at[0] = fmaf(qs, lg, at[0]);
        as[0] = fmaf(qt, lg, as[0]);
        if constexpr (KERNEL == BTK_KERNEL_POT_GRAD) {
            const float rc = __builtin_amdgcn_rcpf(R);
            const float g = qs * rc, h = qt * rc;
            at[1] = fmaf(g, dx, at[1]);
            at[2] = fmaf(g, dy, at[2]);
            as[1] = fmaf(-h, dx, as[1]);
            as[2] = fmaf(-h, dy, as[2]);
        }
    }
}

// Coordinates of the padding lanes of a partial tile: far from every point
// (|d| ~ 1e15, so R stays finite in FP32) and with zero charge, so they add
// nothing to the targets; their own sums are dropped.
constexpr double kFar = 1.0e15;
constexpr int kMaxWaves = 8;

// blockDim.x = 64 * W.  Item: targets [first, first + W*64*TPT) of panel T
// and the owned mutual entries of [e0, e1) (entries with mmeta.z == 0 are
// not owned).  A round takes up to kTiles tiles of 64 points of S into LDS,
// each stored twice (slots j and j + 64 of its block) so that step k of
// lane i reads slot i + k without wrapping; every wavefront runs its 64
// steps per tile, parks the tile's per-point sums in LDS, and after the
// round's barrier the workgroup adds the W wavefronts' sums of each point
// and writes them (one atomic add per point and component).
constexpr int kTiles = 4;

// WMAX: the largest wavefront count the launch may use (4 or 8); it sizes the
// cross-wavefront reduction buffer, so a launch of 4 wavefronts per workgroup
// takes half the LDS and runs more workgroups per CU.
template <int DIM, int KERNEL, int TPT, int WMAX>
__global__ void __launch_bounds__(kLanes * WMAX)
    k_mutual(TaskView tk, EntryView mv, const int *__restrict__ start, const int *__restrict__ count,
             const int *__restrict__ perm, const double *__restrict__ px, const double *__restrict__ py,
             const double *__restrict__ pz, const double *__restrict__ q, KParams kp, double *__restrict__ out,
             size_t comp_stride) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    constexpr int SLOTS = kTiles * 2 * kLanes;
    __shared__ double lx[SLOTS], ly[SLOTS], lz[DIM == 3 ? SLOTS : 1];
    __shared__ float lq[SLOTS];
    __shared__ float red[C][WMAX][kTiles * kLanes];
    const int tid = threadIdx.x;
    const int lane = tid % kLanes;
    const int w = tid / kLanes;
    const int W = blockDim.x / kLanes;
    const int item = xcd_contiguous(blockIdx.x, gridDim.x);
    const int T = tk.panel[item];
    const int first = tk.first[item] + w * kLanes * TPT;  // this wavefront's first target
    const int e0 = tk.e0[item], e1 = tk.e1[item];
    const int nT = count[T], t0 = start[T];
    const bool active = first < nT;  // wavefronts past the leaf's last target only join the reductions
    const float eps2 = (float)kp.eps2;
    double t[TPT][3];
    float qt[TPT];
    double acc[TPT][C];
#pragma unroll
    for (int k = 0; k < TPT; k++) {
        const int ti = first + lane + kLanes * k;
        const int tc = max(0, min(ti, nT - 1));
        t[k][0] = px[t0 + tc];
        t[k][1] = py[t0 + tc];
        t[k][2] = DIM == 3 ? pz[t0 + tc] : 0.0;
        qt[k] = ti < nT ? (float)q[t0 + tc] : 0.0f;  // padding targets act on no source
#pragma unroll
        for (int c = 0; c < C; c++) acc[k][c] = 0.0;
    }
    for (int e = e0; e < e1; e++) {
        const int4 m = mv.meta[e];
        const int nS = m.z, s0 = m.y;
        if (nS == 0) continue;
        for (int r0 = 0; r0 < nS; r0 += kTiles * kLanes) {
            const int ntile = min(kTiles, (nS - r0 + kLanes - 1) / kLanes);
            for (int sl = tid; sl < ntile * 2 * kLanes; sl += blockDim.x) {
                const int j = r0 + (sl / (2 * kLanes)) * kLanes + (sl % kLanes);
                const bool ok = j < nS;
                lx[sl] = ok ? px[s0 + j] : kFar;
                ly[sl] = ok ? py[s0 + j] : kFar;
                if constexpr (DIM == 3) lz[sl] = ok ? pz[s0 + j] : kFar;
                lq[sl] = ok ? (float)q[s0 + j] : 0.0f;
            }
            __syncthreads();
            for (int b = 0; b < ntile; b++) {
                float as[C];
#pragma unroll
                for (int c = 0; c < C; c++) as[c] = 0.0f;
                if (active) {
                    float at[TPT][C];
#pragma unroll
                    for (int k = 0; k < TPT; k++)
#pragma unroll
                        for (int c = 0; c < C; c++) at[k][c] = 0.0f;
                    const int base = b * 2 * kLanes + lane;
#pragma unroll 8
                    for (int step = 0; step < kLanes; step++) {
                        const int sl = base + step;
                        const double sx = lx[sl], sy = ly[sl], sz = DIM == 3 ? lz[sl] : 0.0;
                        const float qs = lq[sl];
#pragma unroll
                        for (int k = 0; k < TPT; k++)
                            eval_both<DIM, KERNEL>(t[k], qt[k], sx, sy, sz, qs, eps2, at[k], as);
                        // the sums of point (lane + step + 1) are in lane + 1
#pragma unroll
                        for (int c = 0; c < C; c++) as[c] = rol1(as[c]);
                    }
#pragma unroll
                    for (int k = 0; k < TPT; k++)
#pragma unroll
                        for (int c = 0; c < C; c++) acc[k][c] += (double)at[k][c];
                }
                // after 64 rotations lane i holds the sums of point r0 + 64 b + i
#pragma unroll
                for (int c = 0; c < C; c++) red[c][w][b * kLanes + lane] = as[c];
            }
            __syncthreads();
            for (int i = tid; i < C * ntile * kLanes; i += blockDim.x) {
                const int c = i / (ntile * kLanes), jj = i % (ntile * kLanes);
                const int j = r0 + jj;
                if (j < nS) {
                    float sum = 0.0f;
                    for (int v = 0; v < W; v++) sum += red[c][v][jj];
                    unsafeAtomicAdd(out + c * comp_stride + (size_t)(perm ? perm[s0 + j] : s0 + j),
                                    comp_scale_f32<DIM, KERNEL>(c, kp.gfc) * (double)sum);
                }
            }
            // the next round's tile and red writes wait for this barrier
            __syncthreads();
        }
    }
    if (!active) return;
#pragma unroll
    for (int k = 0; k < TPT; k++) {
        const int ti = first + lane + kLanes * k;
        if (ti < nT) {
            const size_t o = (size_t)(perm ? perm[t0 + ti] : t0 + ti);
#pragma unroll
            for (int c = 0; c < C; c++)
                unsafeAtomicAdd(out + c * comp_stride + o, comp_scale_f32<DIM, KERNEL>(c, kp.gfc) * acc[k][c]);
        }
    }
}

}  // namespace

void build_mutual(const List &l, const Tree &t, int tpt, int4 *meta, Mutual &mu, hipStream_t s) {
    mu.n = 0;
    const int pct = t.pc;
    if (l.count == 0 || pct == 0) return;
    // near entries are [0, goff[pct]); the far ones are never mutual
    int *h = reinterpret_cast<int *>(host_scratch().v);
    BTK_HIP(hipMemcpyAsync(h, l.goff.p + pct, sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    const int nn = h[0];
    if (nn == 0) return;
    mu.mmeta.resize(nn);
    k_mutual_mark<<<nblk(nn), TB, 0, s>>>(nn, pct, l.goff.p, l.gsrc.p, l.gtype.p, t.start.p, t.count.p, meta,
                                          mu.mmeta.p);
    BTK_LAUNCH_CHECK();
    // W wavefronts cover an average leaf (at most 8)
    const double avg = t.n_leaves ? (double)t.n / t.n_leaves : 1.0;
    mu.tpt = tpt;
    mu.waves = std::max(1, std::min(kMaxWaves, (int)std::ceil(avg / (kLanes * tpt))));
    const int per = kLanes * tpt * mu.waves;
    const long long base = (long long)t.n / per + t.n_leaves;
    const int nsplit = choose_split(base, 8192);  // about 8 workgroups per CU
    build_tasks(l, t, t, 0, 1, per, nsplit, mu.tasks, s, mu.mmeta.p);
    mu.n = mu.tasks.n;
}

template <int DIM, int KERNEL>
void launch_mutual(const Mutual &mu, const Tree &t, const double *q_sorted, const KParams &kp, double *out,
                   bool sorted_out, hipStream_t s) {
    if (mu.n <= 0) return;
    const size_t stride = (size_t)t.n;
    const TaskView tk{mu.tasks.panel.p, mu.tasks.first.p, mu.tasks.e0.p, mu.tasks.e1.p};
    const EntryView mv{mu.mmeta.p};
    const int *perm = sorted_out ? nullptr : t.perm.p;
#define BTK_MU(TPT_, WMAX_)                                                                                   \
    k_mutual<DIM, KERNEL, TPT_, WMAX_><<<mu.n, kLanes * mu.waves, 0, s>>>(tk, mv, t.start.p, t.count.p, perm, \
                                                                           t.sx.p, t.sy.p, t.sz.p, q_sorted,   \
                                                                           kp, out, stride)
    BTK_REQUIRE(mu.waves >= 1 && mu.waves <= kMaxWaves, "mutual pairs: 1 to 8 wavefronts per workgroup");
    if (mu.tpt == 2) {
        if (mu.waves <= 4)
            BTK_MU(2, 4);
        else
            BTK_MU(2, 8);
    } else {
        if (mu.waves <= 4)
            BTK_MU(1, 4);
        else
            BTK_MU(1, 8);
    }
#undef BTK_MU
    BTK_LAUNCH_CHECK();
}

template void launch_mutual<2, BTK_KERNEL_POT>(const Mutual &, const Tree &, const double *, const KParams &, double *,
                                               bool, hipStream_t);
template void launch_mutual<2, BTK_KERNEL_POT_GRAD>(const Mutual &, const Tree &, const double *, const KParams &,
                                                    double *, bool, hipStream_t);
template void launch_mutual<3, BTK_KERNEL_POT>(const Mutual &, const Tree &, const double *, const KParams &, double *,
                                               bool, hipStream_t);
template void launch_mutual<3, BTK_KERNEL_REF_VEL3D>(const Mutual &, const Tree &, const double *, const KParams &,
                                                     double *, bool, hipStream_t);
template void launch_mutual<3, BTK_KERNEL_POT_GRAD>(const Mutual &, const Tree &, const double *, const KParams &,
                                                    double *, bool, hipStream_t);

}  // namespace btk
