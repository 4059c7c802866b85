// MFMA proxy products: many charge vectors over one geometry, and CC pairs
// sharing a kernel matrix (neither has a reference counterpart; the
// reference solves one charge vector per run, pair by pair).
//
// Many charge vectors.
// Tree, list, basis matrices and kernel evaluations are shared by all K
// vectors, so the interactions become products G Q of the kernel tile G
// (targets x sources) with the charge tile Q (sources x K):
//   a workgroup of KC/16 wavefronts owns 64 targets of one panel and KC
//   charge vectors (KC = 32, 64 or 128, btk_tuning.multi_vectors); wave w
//   takes targets 32*(w&1) .. +32 and vectors 32*(w>>1) .. +32;
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
inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }
// threads of a many-vector workgroup of KC charge vectors: two wavefronts
// (64 targets) per 32 vectors, KC / 16 wavefronts in all
constexpr int mrhs_threads(int kc) { return 4 * kc; }

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

// Occupancy asked of the compiler: 5 wavefronts per SIMD for one component
// (96 VGPRs, no spills; 4 by its own choice), 2 for three or four (1).
template <int DIM, int KERNEL, bool FAR, bool EXCL, int KC>
__global__ void __launch_bounds__(mrhs_threads(KC), (KernelTraits<DIM, KERNEL>::C == 1 ? 5 : 2))
    k_interact_mrhs(TaskView tk, GroupView gv, TgtView tv, SrcView sv, MView mv, int n1, NodeTables tabs,
                    KParams kp, int M, int pct, int atomic, int same_tree) {
    constexpr int C = KernelTraits<DIM, KERNEL>::C;
    constexpr int NT = mrhs_threads(KC);
    static_assert(KC == 32 || KC == 64 || KC == 128, "32, 64 or 128 charge vectors per workgroup");
    static_assert(NT == 64 * 2 * (KC / 32), "wave w: targets 32 * (w & 1), vectors 32 * (w >> 1)");
    __shared__ double sc[3][64];
    __shared__ float scr[3][64];  // relative to the target box centre (separated boxes)
    __shared__ float qt[64][KC + 1];
    __shared__ double cn[3][kMaxN1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int item = xcd_contiguous(blockIdx.x, gridDim.x);
    const int T = tk.panel[item];
    const int first = tk.first[item];
    const int e0 = tk.e0[item], e1 = tk.e1[item];
    const int rhs0 = blockIdx.y * KC;
    if (rhs0 >= mv.K) return;  // padding vectors only (the whole workgroup)
    const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
    const int ntgt = FAR ? P : tv.count[T];
    const int ts0 = FAR ? 0 : tv.start[T];
    // a wavefront whose 32 vectors or 32 targets are all padding keeps to the
    // barriers only
    const bool idle = rhs0 + 32 * (w >> 1) >= mv.K || first + 32 * (w & 1) >= ntgt;
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
    // far pairs, and near pairs of separated boxes, take FP32 coordinates
    // relative to the target box centre (the error bound of the one-vector kernels)
    const double ctr[3] = {0.5 * (tv.lo0[T] + tv.hi0[T]), 0.5 * (tv.lo1[T] + tv.hi1[T]),
                           DIM == 3 ? 0.5 * (tv.lo2[T] + tv.hi2[T]) : 0.0};
    const float tr[3] = {(float)(t[0] - ctr[0]), (float)(t[1] - ctr[1]), (float)(t[2] - ctr[2])};
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
        const bool rel = FAR || separated_for_rel<DIM>(T, S, tv, sv);
        const bool check = EXCL && !rel && (!same_tree || FAR || cluster || S == T);
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
                if (rel) {
                    scr[0][tid] = (float)(x - ctr[0]);
                    scr[1][tid] = (float)(y - ctr[1]);
                    scr[2][tid] = (float)(z - ctr[2]);
                } else {
                    sc[0][tid] = x;
                    sc[1][tid] = y;
                    sc[2][tid] = z;
                }
            }
            const float *qsrc = cluster ? mv.w + ((size_t)S * P + t0) * mv.kp : mv.q + (size_t)(ss + t0) * mv.kp;
            // columns past the padded vector count (kp is a multiple of 64,
            // the last block of 128 may reach past it) read as zero
            for (int k = tid; k < 64 * KC; k += NT) {
                const int s = k / KC, r = k % KC;
                qt[s][r] = (s < nt && rhs0 + r < mv.kp) ? qsrc[(size_t)s * mv.kp + rhs0 + r] : 0.0f;
            }
            __syncthreads();
            f32x16 d[C];
#pragma unroll
            for (int c = 0; c < C; c++) d[c] = (f32x16){0};
            const int nsteps = (nt + 1) / 2;  // a partial last tile runs only its sources' steps
            auto steps = [&](auto relc) {
#pragma unroll 4
            for (int step = 0; step < nsteps; step++) {
                const int s = 2 * step + par;
                float dx, dy, dz;
                if constexpr (decltype(relc)::value) {
                    dx = tr[0] - scr[0][s];
                    dy = tr[1] - scr[1][s];
                    dz = DIM == 3 ? tr[2] - scr[2][s] : 0.0f;
                } else {
                    dx = (float)(t[0] - sc[0][s]);
                    dy = (float)(t[1] - sc[1][s]);
                    dz = DIM == 3 ? (float)(t[2] - sc[2][s]) : 0.0f;
                }
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
            };
            if (!idle) {
                if (rel)
                    steps(std::true_type{});
                else
                    steps(std::false_type{});
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

// ---------------------------------------------------------------------------
// CC pairs sharing a kernel matrix (CCGemm, btk_interact.hpp).
// ---------------------------------------------------------------------------
struct LevelBoxes {
    const double *lo0, *lo1, *lo2, *hi0, *hi1, *hi2;
    const int *level;
};
LevelBoxes level_boxes(const Tree &t) {
    return LevelBoxes{t.lo[0].p, t.lo[1].p, t.lo[2].p, t.hi[0].p, t.hi[1].p, t.hi[2].p, t.level.p};
}

constexpr unsigned long long kNoKey = ~0ull;
constexpr int kCcTgt = 128;  // target proxies per work item (32 per wavefront)
constexpr int kCcCols = 32;  // pairs per work item

// Target panel of entry e of a group whose offsets are goff_g[0..pct].
__device__ __forceinline__ int group_panel(const int *__restrict__ goff_g, int pct, int e) {
    int lo = 0, hi = pct;  // goff_g[lo] <= e < goff_g[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (goff_g[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Key of entry e0 + i: problem (forests: the target panel's root), levels,
// and centre offset in half boxes of the finer level, ob bits per axis; the
// key's width is pbits + 12 + 3 ob <= 63 bits.  kNoKey for other types, taken
// entries, and offsets off the lattice or out of range.
struct KeyLayout {
    int nprob, pbits, ob;
};

__global__ void k_cc_keys(int n, int e0, const int *__restrict__ goff_g, int pct, const int4 *__restrict__ meta,
                          LevelBoxes T, LevelBoxes S, const int *__restrict__ tparent, KeyLayout kl,
                          unsigned long long *__restrict__ keys, int *__restrict__ vals) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int e = e0 + i;
    vals[i] = e;
    const int4 m = meta[e];
    unsigned long long key = kNoKey;
    if ((m.w & 3) == 3 && m.z > 0) {
        const int t = group_panel(goff_g, pct, e), s = m.x;
        const int lt = T.level[t], ls = S.level[s];
        int root = t;
        while (root >= kl.nprob) root = tparent[root];
        const long long half = 1ll << (kl.ob - 1);
        bool ok = lt < 63 && ls < 63;  // (a key of all ones would tie with kNoKey)
        bool apart = false;  // separated boxes: no zero-distance proxy pair (k_cc_gemm has no such test)
        unsigned long long k = ((((unsigned long long)root << 6) | (unsigned long long)lt) << 6 | ls) << (3 * kl.ob);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double tl = sel3(d, T.lo0[t], T.lo1[t], T.lo2[t]), th = sel3(d, T.hi0[t], T.hi1[t], T.hi2[t]);
            const double sl = sel3(d, S.lo0[s], S.lo1[s], S.lo2[s]), sh = sel3(d, S.hi0[s], S.hi1[s], S.hi2[s]);
            apart = apart || sl > th || tl > sh;
            const double h = 0.5 * fmin(th - tl, sh - sl);
            const double off = 0.5 * (sl + sh) - 0.5 * (tl + th);
            long long q = 0;
            if (h > 0) {
                const double r = off / h;
                q = llrint(r);
                ok = ok && fabs(r - (double)q) <= 1e-9 && q > -half && q < half;
            } else {
                ok = ok && off == 0.0;
            }
            k |= (unsigned long long)(q + half) << (kl.ob * (2 - d));
        }
        if (ok && apart) key = k;
    }
    keys[i] = key;
}

__global__ void k_cc_heads(int n, const unsigned long long *__restrict__ keys, int *__restrict__ head) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    head[i] = (keys[i] != kNoKey && (i == 0 || keys[i] != keys[i - 1])) ? 1 : 0;
}

// Start of each key (gid: inclusive count of heads) and the number of keyed
// entries (they sort before the unkeyed ones).
__global__ void k_cc_starts(int n, const unsigned long long *__restrict__ keys, const int *__restrict__ head,
                            const int *__restrict__ gid, int *__restrict__ gstart, int *__restrict__ nvalid) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    if (head[i]) gstart[gid[i] - 1] = i;
    if (keys[i] != kNoKey && (i + 1 == n || keys[i + 1] == kNoKey)) *nvalid = i + 1;
}

// A key fixes the geometry when every box of a level has one size, as in a
// tree of one problem; each keyed entry still checks its boxes' extents
// against the key's first pair and spoils the key (bad) on a mismatch.
__global__ void k_cc_verify(int nv, const int *__restrict__ vals, const int *__restrict__ goff_g, int pct,
                            const int4 *__restrict__ meta, const int *__restrict__ gid,
                            const int *__restrict__ gstart, LevelBoxes T, LevelBoxes S, int *__restrict__ bad) {
    const int k = blockIdx.x * TB + threadIdx.x;
    if (k >= nv) return;
    const int g = gid[k] - 1;
    const int e = vals[k], e0 = vals[gstart[g]];
    const int t = group_panel(goff_g, pct, e), t0 = group_panel(goff_g, pct, e0);
    const int s = meta[e].x, s0 = meta[e0].x;
    bool same = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double a = sel3(d, T.hi0[t], T.hi1[t], T.hi2[t]) - sel3(d, T.lo0[t], T.lo1[t], T.lo2[t]);
        const double a0 = sel3(d, T.hi0[t0], T.hi1[t0], T.hi2[t0]) - sel3(d, T.lo0[t0], T.lo1[t0], T.lo2[t0]);
        const double b = sel3(d, S.hi0[s], S.hi1[s], S.hi2[s]) - sel3(d, S.lo0[s], S.lo1[s], S.lo2[s]);
        const double b0 = sel3(d, S.hi0[s0], S.hi1[s0], S.hi2[s0]) - sel3(d, S.lo0[s0], S.lo1[s0], S.lo2[s0]);
        same = same && fabs(a - a0) <= 1e-9 * fabs(a0) && fabs(b - b0) <= 1e-9 * fabs(b0);
    }
    if (!same) bad[g] = 1;
}

__global__ void k_cc_count(int G, int *__restrict__ gstart, const int *__restrict__ nvalid, const int *__restrict__ bad,
                           int min_group, int *__restrict__ nitems, unsigned long long *__restrict__ taken) {
    const int g = blockIdx.x * TB + threadIdx.x;
    if (g >= G) return;
    const int end = g + 1 < G ? gstart[g + 1] : *nvalid;
    const int size = end - gstart[g];
    const bool sel = size >= min_group && !bad[g];
    nitems[g] = sel ? (size + kCcCols - 1) / kCcCols : 0;
    if (sel) atomicAdd(taken, (unsigned long long)size);
}

__global__ void k_cc_fill(int G, const int *__restrict__ gstart, const int *__restrict__ nvalid,
                          const int *__restrict__ nitems, const int *__restrict__ ibase,
                          int4 *__restrict__ items) {
    const int g = blockIdx.x;
    const int ni = nitems[g];
    if (ni == 0) return;
    const int end = g + 1 < G ? gstart[g + 1] : *nvalid;
    const int size = end - gstart[g];
    for (int it = threadIdx.x; it < ni; it += blockDim.x)
        items[ibase[g] + it] = make_int4(gstart[g], size, 0, kCcCols * it);
}

// Work items in the order of the target cluster of their first pair (stable:
// keys in key order within one cluster).  A key holds each target cluster at
// most once, so a cluster's proxy sums receive one item of each of its keys;
// in key order those items lie a whole level's clusters apart, in this order
// they run close together and the sums' cache lines are still in L2 / MALL
// when the next item adds into them.
__global__ void k_cc_item_keys(int n, const int4 *__restrict__ items, const int *__restrict__ entT,
                               int *__restrict__ key, int *__restrict__ val) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int4 it = items[i];
    key[i] = entT[it.x + it.w];
    val[i] = i;
}

__global__ void k_cc_item_gather(int n, const int *__restrict__ order, const int4 *__restrict__ in,
                                 int4 *__restrict__ out) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i < n) out[i] = in[order[i]];
}

// Panels of every keyed entry in key order; the entries of the selected keys
// leave the regular far kernel (source count zero in meta).
__global__ void k_cc_entries(int nv, const int *__restrict__ vals, const int *__restrict__ goff_g, int pct,
                             const int *__restrict__ gid, const int *__restrict__ nitems, int4 *__restrict__ meta,
                             int *__restrict__ entT, int *__restrict__ entS) {
    const int k = blockIdx.x * TB + threadIdx.x;
    if (k >= nv) return;
    const int e = vals[k];
    entS[k] = meta[e].x;
    entT[k] = group_panel(goff_g, pct, e);
    if (nitems[gid[k] - 1] > 0) meta[e].z = 0;
}

// Keyed entry k sorted by its target panel when an item covers it (its key
// has work items), after every panel otherwise.
__global__ void k_cc_taken_keys(int nv, const int *__restrict__ gid, const int *__restrict__ nitems,
                                const int *__restrict__ entT, int pct, int *__restrict__ key, int *__restrict__ val) {
    const int k = blockIdx.x * TB + threadIdx.x;
    if (k >= nv) return;
    key[k] = nitems[gid[k] - 1] > 0 ? entT[k] : pct;
    val[k] = k;
}

// off[p] = first position of sorted with sorted[.] >= p, p = 0 .. pct.
__global__ void k_csr_offsets(int n, const int *__restrict__ sorted, int pct, int *__restrict__ off) {
    const int p = blockIdx.x * TB + threadIdx.x;
    if (p > pct) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sorted[mid] < p) lo = mid + 1;
        else hi = mid;
    }
    off[p] = lo;
}

// Target panel p's potentials of components 0..C-1 += the staged products of
// the keyed entries tk[off[p] .. off[p+1]), summed in FP64 in that order
// (per chunk of kRedChunk entries); one workgroup per panel.  The entries'
// indices sit in LDS, and each thread issues the loads of 8 entries together
// before adding them in order.
constexpr int kRedChunk = 1024;
__global__ void __launch_bounds__(TB)
    k_cc_reduce(int C, int P, int Pst, const int *__restrict__ off, const int *__restrict__ tk,
                const float *__restrict__ stage, double *__restrict__ out, size_t comp_stride) {
    __shared__ int sk[kRedChunk];
    const int p = blockIdx.x;
    const int e0 = off[p], e1 = off[p + 1];
    if (e0 == e1) return;
    const size_t rs = (size_t)C * Pst;  // one entry's staged floats
    for (int c0 = e0; c0 < e1; c0 += kRedChunk) {
        const int ne = min(kRedChunk, e1 - c0);
        __syncthreads();
        for (int e = threadIdx.x; e < ne; e += TB) sk[e] = tk[c0 + e];
        __syncthreads();
        for (int i = threadIdx.x; i < C * P; i += TB) {
            const float *col = stage + (i / P) * Pst + i % P;
            double sum = 0.0;
            int e = 0;
            for (; e + 8 <= ne; e += 8) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = col[(size_t)sk[e + u] * rs];
#pragma unroll
                for (int u = 0; u < 8; u++) sum += (double)v[u];
            }
            for (; e < ne; e++) sum += (double)col[(size_t)sk[e] * rs];
            out[(i / P) * comp_stride + (size_t)p * P + i % P] += sum;
        }
    }
}

// Split-FP16 operands (Ootomo and Yokota, "Recovering single precision
// accuracy from Tensor Cores while surpassing the FP32 theoretical peak
// performance", IJHPCA 36(4), 2022): an FP32 value x travels as hi = fp16(x)
// and lo = fp16(x - hi), both rounded to nearest, and a product as
// hi_a hi_b + lo_a hi_b + hi_a lo_b; the dropped lo_a lo_b is below 2^-22 of
// the product, and every fp16 product is exact in the FP32 accumulators.
// Power-of-two scalings (exact) keep the operands below 2^14, inside the
// fp16 range (largest finite value 65504) with room for rounding.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
constexpr int kCcExp = 14;

// m1 = -1: the residuals x - hi (exact in FP32) are then fmas of an fp16
// operand, one v_fma_mix_f32 each instead of a conversion and a subtraction;
// the caller keeps m1 opaque, in a VGPR.
__device__ __forceinline__ void split2(float a, float b, float m1, unsigned &hi, unsigned &lo) {
    const half2v h = __builtin_convertvector((float2v){a, b}, half2v);
    const half2v l = __builtin_convertvector((float2v){__builtin_fmaf((float)h.x, m1, a), __builtin_fmaf((float)h.y, m1, b)},
                                             half2v);
    hi = __builtin_bit_cast(unsigned, h);
    lo = __builtin_bit_cast(unsigned, l);
}

__device__ __forceinline__ half8 as_half8(uint4 v) { return __builtin_bit_cast(half8, v); }

// A workgroup barrier that orders LDS accesses only.  __syncthreads() also
// makes the workgroup's global writes visible, which waits for every
// outstanding global store and atomic; the CC products' atomic adds go to
// sums no thread of the launch reads, so their barriers need not wait.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Source proxy weights of one degree as split-FP16 rows, one wavefront per
// source panel: the row is scaled by 2^e (e = wexp[panel]) so that its
// largest magnitude lies in [2^13, 2^14), and padded with zeros to the row
// stride Ps (a multiple of 64).
__global__ void k_weights_f16(int pc, int P, int Ps, const double *__restrict__ W, unsigned short *__restrict__ Wh,
                              unsigned short *__restrict__ Wl, int *__restrict__ wexp) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    if (c >= pc) return;
    const double *w = W + (size_t)c * P;
    double m = 0.0;
    for (int j = lane; j < P; j += 64) m = fmax(m, fabs(w[j]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    int e = 0;
    if (m > 0.0) {
        int ex;
        (void)frexp(m, &ex);  // 2^(ex-1) <= m < 2^ex
        e = kCcExp - ex;
    }
    if (lane == 0) wexp[c] = e;
    for (int j = 2 * lane; j < Ps; j += 128) {
        const float a = j < P ? (float)ldexp(w[j], e) : 0.0f;
        const float b = j + 1 < P ? (float)ldexp(w[j + 1], e) : 0.0f;
        unsigned hi, lo;
        split2(a, b, -1.0f, hi, lo);
        *reinterpret_cast<unsigned *>(Wh + (size_t)c * Ps + j) = hi;
        *reinterpret_cast<unsigned *>(Wl + (size_t)c * Ps + j) = lo;
    }
}

// One work item: 128 target proxies x up to 32 pairs of one key, components
// C0 .. C0+NC-1 of the kernel (every component in one launch: each kernel
// value feeds NC products).  Wave w takes targets 32*w .. +32 and all 32
// pairs and forms the transposed product, pairs x targets, on fp16 matrix
// cores (v_mfma_f32_32x32x16_f16, 16 sources per step, split-FP16 operands,
// three products per step and component): the A operand holds the pairs'
// proxy weights (split rows from k_weights_f16, staged in LDS one 64-source
// tile at a time, the next tile's loads in flight), the B operand the kernel
// of the key's first pair (FP32 coordinates relative to its target box
// centre, as the far kernel's TileRel, scaled by a power of two), lane l
// evaluating it for target l%32 and sources 8*(l/32) .. +8 of the step.  The
// fp16 products overlap the kernel evaluations on the vector ALU.
// The matrix cores sum a pair's sources in FP32 (3 * Pc / 16 accumulations,
// 138 at degree 8); each pair's sums, unscaled, are stored to the pair's own
// FP32 rows of `stage` (Pst floats per component, 128-byte aligned), and
// k_cc_reduce adds the pairs of each target cluster in FP64.  Each
// accumulator row is one pair, its lanes 32 consecutive target proxies, so
// the stores are coalesced.  (FP64 atomic adds into the clusters' sums, one
// per pair, proxy and component, ran at about 4e11 per second: 19 ms of the
// 16M Plummer solve, where the stores run under the products and the
// reduction takes 7 ms.)
struct CcWeights {
    const unsigned short *hi, *lo;  // [source panel][Ps] split-FP16 weights of one degree
    const int *exp;                 // [source panel] their power-of-two scale
    int stride;                     // Ps
};

template <int KERNEL, int C0, int NC>
__global__ void __launch_bounds__(TB, 2)
    k_cc_gemm(const int4 *__restrict__ items, const int *__restrict__ entT, const int *__restrict__ entS,
              LevelBoxes TB_, LevelBoxes SB, CcWeights wq, int n1, NodeTables tabs, KParams kp,
              float *__restrict__ stage, int Pst) {
    constexpr int C = KernelTraits<3, KERNEL>::C;
    static_assert(C0 + NC <= C, "component range");
    constexpr int kMaxP = kMaxN1 * kMaxN1 * kMaxN1;
    constexpr int kMaxPs = (kMaxP + 63) / 64 * 64;
    __shared__ __align__(16) float sc[3][kMaxPs];  // every source proxy of the key's first pair, relative to ctr
    __shared__ float tc[3][kMaxP];                 // every target proxy of that pair, relative to ctr
    __shared__ uint4 qs[2][kCcCols][9];  // hi, lo weights of the tile: 8 x 8 sources per pair, 1 of padding
    __shared__ double cn[2][3][kMaxN1];  // source, target box nodes
    __shared__ int sS[kCcCols], sT[kCcCols], sE[kCcCols];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int4 it = items[xcd_contiguous(blockIdx.x, gridDim.x)];
    const int gs = it.x, gn = it.y, p0 = it.w;
    const int T0 = entT[gs], S0 = entS[gs];
    const int P = n1 * n1 * n1;
    const int Pc = (P + 15) & ~15;  // sources evaluated: whole 16-source steps
    const int Ps = wq.stride;       // >= Pc, a multiple of 64
    const double ctr[3] = {0.5 * (TB_.lo0[T0] + TB_.hi0[T0]), 0.5 * (TB_.lo1[T0] + TB_.hi1[T0]),
                           0.5 * (TB_.lo2[T0] + TB_.hi2[T0])};
    const int col = lane & 31;
    const int par = lane >> 5;
    if (tid < kCcCols) {
        const bool v = p0 + tid < gn;
        const int S = v ? entS[gs + p0 + tid] : -1;
        sS[tid] = S;
        sT[tid] = v ? entT[gs + p0 + tid] : -1;
        sE[tid] = wq.exp[S >= 0 ? S : S0];
    }
    if (tid < 3 * n1) {
        const int d = tid / n1, b = tid % n1;
        cn[0][d][b] = proxy_node(tabs.b[b], sel3(d, SB.lo0[S0], SB.lo1[S0], SB.lo2[S0]),
                                 sel3(d, SB.hi0[S0], SB.hi1[S0], SB.hi2[S0]));
        cn[1][d][b] = proxy_node(tabs.b[b], sel3(d, TB_.lo0[T0], TB_.lo1[T0], TB_.lo2[T0]),
                                 sel3(d, TB_.hi0[T0], TB_.hi1[T0], TB_.hi2[T0]));
    }
    // The boxes of a key are apart: every proxy pair is at least their gap g
    // apart, and with 2^(eg-1) <= g the scaled kernel values 2^(kCcExp+eg-1)/r
    // and 2^(kCcExp+2eg-2) |d|/r^3 stay below 2^kCcExp.
    double g2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double tl = sel3(d, TB_.lo0[T0], TB_.lo1[T0], TB_.lo2[T0]), th = sel3(d, TB_.hi0[T0], TB_.hi1[T0], TB_.hi2[T0]);
        const double sl = sel3(d, SB.lo0[S0], SB.lo1[S0], SB.lo2[S0]), sh = sel3(d, SB.hi0[S0], SB.hi1[S0], SB.hi2[S0]);
        const double gap = fmax(0.0, fmax(sl - th, tl - sh));
        g2 = fma(gap, gap, g2);
    }
    int eg;
    (void)frexp(sqrt(g2), &eg);
    const float fpot = ldexpf(1.0f, kCcExp + eg - 1), fgrad = ldexpf(1.0f, kCcExp + 2 * (eg - 1));
    // The coordinates carry the potential's scale: with d' = d / fpot,
    // 1/|d'| = fpot/|d| and |d'|/|d'|^3 = 2^kCcExp * fgrad |d|/|d|^3 (powers of
    // two, exact), so the potential needs no product and the gradient one.
    const double sig = ldexp(1.0, -(kCcExp + eg - 1));
    __syncthreads();
    // source and target proxy coordinates relative to the target box
    // centre, scaled by sig, once per item (proxy_point's order); the
    // padding sources P .. Pc (zero weights) sit at the last proxy
    for (int j = tid; j < Pc; j += TB) {
        const int jj = min(j, P - 1);
        const int a = jj / (n1 * n1), b = (jj / n1) % n1, c = jj % n1;
        sc[0][j] = (float)((cn[0][0][a] - ctr[0]) * sig);
        sc[1][j] = (float)((cn[0][1][b] - ctr[1]) * sig);
        sc[2][j] = (float)((cn[0][2][c] - ctr[2]) * sig);
        if (j < P) {
            tc[0][j] = (float)((cn[1][0][a] - ctr[0]) * sig);
            tc[1][j] = (float)((cn[1][1][b] - ctr[1]) * sig);
            tc[2][j] = (float)((cn[1][2][c] - ctr[2]) * sig);
        }
    }
    lds_barrier();
    // this thread's share of a weight tile: pair r = tid / 8, sources 8 * (tid % 8) .. +8.
    // A padding pair row reads the key's first weights (its output row is
    // never written; rows of a product do not mix).
    auto load_q = [&](int t0, uint4 &h, uint4 &l) {
        const int S = sS[tid >> 3];
        const size_t o = (size_t)(S >= 0 ? S : S0) * Ps + t0 + 8 * (tid & 7);
        h = *reinterpret_cast<const uint4 *>(wq.hi + o);
        l = *reinterpret_cast<const uint4 *>(wq.lo + o);
    };
    // the item runs every block of kCcTgt targets against every 64-source
    // tile, one sequence of (target block, source tile) steps; the weights
    // of the next step are loaded during the current one
    const int nst = (Pc + 63) / 64;
    const int steps = nst * ((P + kCcTgt - 1) / kCcTgt);
    uint4 nh, nl;
    load_q(0, nh, nl);
    qs[0][tid >> 3][tid & 7] = nh;
    qs[1][tid >> 3][tid & 7] = nl;
    lds_barrier();
    f32x16 d[NC];
    float tr[3] = {0.0f, 0.0f, 0.0f};
    int ti = 0;
    bool idle = false;  // every target of this wavefront is padding (P is not a multiple of 128)
    const float eps2 = (float)(kp.eps2 * sig * sig);
    const float gs14 = ldexpf(1.0f, -kCcExp);
    float m1 = -1.0f;
    asm volatile("" : "+v"(m1));  // split2's -1 in a VGPR, not folded into subtractions
    for (int st = 0; st < steps; st++) {
        const int t0 = (st % nst) * 64;
        if (t0 == 0) {  // a new block of targets
            ti = (st / nst) * kCcTgt + 32 * w + col;
            idle = (st / nst) * kCcTgt + 32 * w >= P;
            const int tj = min(ti, P - 1);
#pragma unroll
            for (int dd = 0; dd < 3; dd++) tr[dd] = tc[dd][tj];
#pragma unroll
            for (int c = 0; c < NC; c++) d[c] = (f32x16){0};
        }
        // the next tile's weights load while this one is evaluated
        if (st + 1 < steps) load_q(((st + 1) % nst) * 64, nh, nl);
        const int ns = min(4, (Pc - t0) >> 4);  // 16-source steps in this tile
        // The boxes of a CC pair are separated, so no kernel value needs the
        // zero-distance test.
        auto step = [&](int k) {
            const int j0 = t0 + 16 * k + 8 * par;
            const half8 ah = as_half8(qs[0][col][2 * k + par]), al = as_half8(qs[1][col][2 * k + par]);
            // the kernel of the lane's target and its 8 sources: the
            // distance part is shared by the components
            float dx[8], dy[8], dz[8], ip[8], ig[8];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const float4 ux = *reinterpret_cast<const float4 *>(&sc[0][j0 + 4 * h]);
                const float4 uy = *reinterpret_cast<const float4 *>(&sc[1][j0 + 4 * h]);
                const float4 uz = *reinterpret_cast<const float4 *>(&sc[2][j0 + 4 * h]);
                const float xs[3][4] = {{ux.x, ux.y, ux.z, ux.w}, {uy.x, uy.y, uy.z, uy.w}, {uz.x, uz.y, uz.z, uz.w}};
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    dx[4 * h + m] = tr[0] - xs[0][m];
                    dy[4 * h + m] = tr[1] - xs[1][m];
                    dz[4 * h + m] = tr[2] - xs[2][m];
                }
            }
#pragma unroll
            for (int m = 0; m < 8; m++) {
                const float R = fmaf(dx[m], dx[m], fmaf(dy[m], dy[m], fmaf(dz[m], dz[m], eps2)));
                const float inv = __builtin_amdgcn_rsqf(R);
                ip[m] = inv;
                ig[m] = inv * inv * (inv * gs14);
            }
            auto value = [&](int c, int m) {
                return c == 0 ? ip[m] : ig[m] * (c == 1 ? dx[m] : c == 2 ? dy[m] : dz[m]);
            };
            // Stage q splits component q's values into the B operands and
            // issues its hi*hi product, component q-1's lo*hi and q-2's
            // hi*lo: the products of one stage run on the matrix cores while
            // the next stage's values are formed on the vector ALU, and the
            // three products into one accumulator are a stage apart.
            uint4 bh[NC], bl[NC];
#pragma unroll
            for (int q = 0; q < NC + 2; q++) {
                if (q < NC) {
#pragma unroll
                    for (int mp = 0; mp < 4; mp++) {
                        unsigned hi, lo;
                        split2(value(C0 + q, 2 * mp), value(C0 + q, 2 * mp + 1), m1, hi, lo);
                        (mp == 0 ? bh[q].x : mp == 1 ? bh[q].y : mp == 2 ? bh[q].z : bh[q].w) = hi;
                        (mp == 0 ? bl[q].x : mp == 1 ? bl[q].y : mp == 2 ? bl[q].z : bl[q].w) = lo;
                    }
                    d[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, as_half8(bh[q]), d[q], 0, 0, 0);
                }
                if (q >= 1 && q - 1 < NC)
                    d[q - 1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, as_half8(bh[q - 1]), d[q - 1], 0, 0, 0);
                if (q >= 2)
                    d[q - 2] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, as_half8(bl[q - 2]), d[q - 2], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        if (!idle) {
            if (ns == 4) {
                step(0);
                step(1);
                step(2);
                step(3);
            } else {
                for (int k = 0; k < ns; k++) step(k);
            }
        }
        lds_barrier();
        // The next tile goes to LDS before this block's stores are issued: a
        // wait for a load also waits for every memory operation issued
        // before it, so the stores are waited for only one tile later.
        if (st + 1 < steps) {
            qs[0][tid >> 3][tid & 7] = nh;
            qs[1][tid >> 3][tid & 7] = nl;
        }
        // D[i][j]: pair row i = 8*(v/4) + 4*(lane/32) + v%4, target column j = lane%32
        if (t0 + 64 >= Pc && ti < P) {  // the block's last source tile: its sums are complete
            double us[NC];
#pragma unroll
            for (int c = 0; c < NC; c++)
                us[c] = comp_scale_f32<3, KERNEL>(C0 + c, kp.gfc) / (double)(C0 + c == 0 ? fpot : fgrad);
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int i = 8 * (v / 4) + 4 * par + (v % 4);
                if (sT[i] < 0) continue;
                const int e = sE[i];
                float *row = stage + (size_t)(gs + p0 + i) * C * Pst + ti;
#pragma unroll
                for (int c = 0; c < NC; c++) row[(size_t)(C0 + c) * Pst] = (float)ldexp((double)d[c][v] * us[c], -e);
            }
        }
        lds_barrier();
    }
}

template <int KERNEL, int C0, int NC>
void launch_cc_part(const CCGemm &cg, const Tree &tt, const Tree &ts, const CcWeights &wq, int n1,
                    const NodeTables &tabs, const KParams &kp, float *stage, int Pst, hipStream_t s) {
    k_cc_gemm<KERNEL, C0, NC><<<cg.n_items, TB, 0, s>>>(cg.items.p, cg.ent_T.p, cg.ent_S.p, level_boxes(tt),
                                                        level_boxes(ts), wq, n1, tabs, kp, stage, Pst);
    BTK_LAUNCH_CHECK();
}

template <int KERNEL>
void launch_cc_gemm(const CCGemm &cg, const Tree &tt, const Tree &ts, const CcWeights &wq, int n1,
                    const NodeTables &tabs, const KParams &kp, float *stage, int Pst, hipStream_t s) {
    if constexpr (KernelTraits<3, KERNEL>::C == 1) {
        launch_cc_part<KERNEL, 0, 1>(cg, tt, ts, wq, n1, tabs, kp, stage, Pst, s);
    } else {
        launch_cc_part<KERNEL, 0, 4>(cg, tt, ts, wq, n1, tabs, kp, stage, Pst, s);
    }
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

template <int DIM, int KERNEL, bool FAR, bool EXCL, int KC>
void launch_mrhs_kc(const TaskView &tk, int n_tasks, const GroupView &gv, const TgtView &tv, const SrcView &sv,
                    const MView &mv, int n1, const NodeTables &tabs, const KParams &kp_, int M, int pct, int atomic,
                    int same_tree, hipStream_t s) {
    const dim3 grid(n_tasks, (mv.kp + KC - 1) / KC);
    k_interact_mrhs<DIM, KERNEL, FAR, EXCL, KC><<<grid, mrhs_threads(KC), 0, s>>>(tk, gv, tv, sv, mv, n1, tabs, kp_,
                                                                                  M, pct, atomic, same_tree);
}

template <int DIM, int KERNEL, bool FAR, bool EXCL>
void launch_mrhs_fe(int kc, const TaskView &tk, int n_tasks, const GroupView &gv, const TgtView &tv,
                    const SrcView &sv, const MView &mv, int n1, const NodeTables &tabs, const KParams &kp_, int M,
                    int pct, int atomic, int same_tree, hipStream_t s) {
    if (kc == 32)
        launch_mrhs_kc<DIM, KERNEL, FAR, EXCL, 32>(tk, n_tasks, gv, tv, sv, mv, n1, tabs, kp_, M, pct, atomic,
                                                   same_tree, s);
    else if (kc == 128)
        launch_mrhs_kc<DIM, KERNEL, FAR, EXCL, 128>(tk, n_tasks, gv, tv, sv, mv, n1, tabs, kp_, M, pct, atomic,
                                                    same_tree, s);
    else
        launch_mrhs_kc<DIM, KERNEL, FAR, EXCL, 64>(tk, n_tasks, gv, tv, sv, mv, n1, tabs, kp_, M, pct, atomic,
                                                   same_tree, s);
}

template <int DIM, int KERNEL>
void launch_mrhs(const TaskView &tk, int n_tasks, const GroupView &gv, const TgtView &tv, const SrcView &sv,
                 const float *qs, const float *wt, double *out, int K, int kp, int n1, const NodeTables &tabs,
                 const KParams &kp_, int M, int pct, int atomic, int same_tree, bool far, bool excl, int kc,
                 hipStream_t s) {
    if (n_tasks <= 0) return;
    const MView mv{qs, wt, out, K, kp};
    if (far) {
        if (excl)
            launch_mrhs_fe<DIM, KERNEL, true, true>(kc, tk, n_tasks, gv, tv, sv, mv, n1, tabs, kp_, M, pct, atomic,
                                                    same_tree, s);
        else
            launch_mrhs_fe<DIM, KERNEL, true, false>(kc, tk, n_tasks, gv, tv, sv, mv, n1, tabs, kp_, M, pct, atomic,
                                                     same_tree, s);
    } else {
        if (excl)
            launch_mrhs_fe<DIM, KERNEL, false, true>(kc, tk, n_tasks, gv, tv, sv, mv, n1, tabs, kp_, M, pct, atomic,
                                                     same_tree, s);
        else
            launch_mrhs_fe<DIM, KERNEL, false, false>(kc, tk, n_tasks, gv, tv, sv, mv, n1, tabs, kp_, M, pct,
                                                      atomic, same_tree, s);
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
    const int kc = multi_vectors();
    const int chunks = (kp + kc - 1) / kc;
    for (int far = 0; far < 2; far++) {
        if (l.counts[2 * far] + l.counts[2 * far + 1] == 0) continue;
        const int n_clusters = tt.pc - tt.n_leaves;
        // base counts wavefronts (kc / 16 per workgroup): aim at 8192 of them
        const long long base = (far ? (long long)n_clusters * ((P + 63) / 64) : (long long)tt.n / 64 + tt.n_leaves) *
                               chunks * (kc / 16);
        const int nsplit = choose_split(base, 8192);
        Tasks tk;
        build_tasks(l, tt, ts, far, P, 64, nsplit, tk, s);
        if (kev) BTK_HIP(hipEventRecord(kev->ev[far ? 2 : 0], s));
        launch_mrhs<DIM, KERNEL>(TaskView{tk.panel.p, tk.first.p, tk.e0.p, tk.e1.p}, tk.n, gv, tv, sv, qs, wt,
                                 far ? Psi : out, K, kp, n1, tabs, kp_, tt.n, tt.pc, nsplit > 1 ? 1 : 0,
                                 (&tt == &ts) ? 1 : 0, far != 0, o.exclude_self != 0, kc, s);
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
    BTK_REQUIRE(kp % 64 == 0 && kp >= K, "charge-vector padding must be a multiple of 64");
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
                                              int, hipStream_t);
template void launch_mrhs<2, BTK_KERNEL_POT_GRAD>(const TaskView &, int, const GroupView &, const TgtView &,
                                                   const SrcView &, const float *, const float *, double *, int, int,
                                                   int, const NodeTables &, const KParams &, int, int, int, int, bool,
                                                   bool, int, hipStream_t);
template void launch_mrhs<3, BTK_KERNEL_POT>(const TaskView &, int, const GroupView &, const TgtView &,
                                              const SrcView &, const float *, const float *, double *, int, int, int,
                                              const NodeTables &, const KParams &, int, int, int, int, bool, bool,
                                              int, hipStream_t);
template void launch_mrhs<3, BTK_KERNEL_POT_GRAD>(const TaskView &, int, const GroupView &, const TgtView &,
                                                   const SrcView &, const float *, const float *, double *, int, int,
                                                   int, const NodeTables &, const KParams &, int, int, int, int, bool,
                                                   bool, int, hipStream_t);


void build_ccgemm(const List &l, const Tree &tt, const Tree &ts, int grp, int min_group, int4 *meta,
                  CCGemm &cg, hipStream_t s) {
    cg.n_items = 0;
    cg.shared = 0;
    // a forest's problems have boxes of their own sizes at every level: the
    // key leads with the problem (one tree for targets and sources)
    if ((tt.nprob > 1 || ts.nprob > 1) && &tt != &ts) return;
    KeyLayout kl{tt.nprob, 0, 16};
    while ((1ll << kl.pbits) < tt.nprob) kl.pbits++;
    kl.ob = std::min(16, (63 - 12 - kl.pbits) / 3);
    if (kl.ob < 6) return;
    const int key_bits = kl.pbits + 12 + 3 * kl.ob;
    const int pct = tt.pc;
    const int *goff_g = l.goff.p + (size_t)grp * pct;
    int *h = reinterpret_cast<int *>(host_scratch().v);
    BTK_HIP(hipMemcpyAsync(h, goff_g, sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipMemcpyAsync(h + 1, goff_g + pct, sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    const int e0 = h[0], n = h[1] - h[0];
    if (n <= 0) return;
    DBuf<unsigned long long> keys(n), keys2(n);
    DBuf<int> vals(n), vals2(n), head(n), gid(n);
    k_cc_keys<<<nblk(n), TB, 0, s>>>(n, e0, goff_g, pct, meta, level_boxes(tt), level_boxes(ts), tt.parent.p, kl,
                                     keys.p, vals.p);
    BTK_LAUNCH_CHECK();
    sort_pairs_u64(keys.p, keys2.p, vals.p, vals2.p, n, key_bits, s);
    k_cc_heads<<<nblk(n), TB, 0, s>>>(n, keys2.p, head.p);
    BTK_LAUNCH_CHECK();
    inclusive_sum_i32(head.p, gid.p, n, s);
    DBuf<int> gstart(n), nvalid(1);
    BTK_HIP(hipMemsetAsync(nvalid.p, 0, sizeof(int), s));
    k_cc_starts<<<nblk(n), TB, 0, s>>>(n, keys2.p, head.p, gid.p, gstart.p, nvalid.p);
    BTK_LAUNCH_CHECK();
    BTK_HIP(hipMemcpyAsync(h, gid.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipMemcpyAsync(h + 1, nvalid.p, sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    const int G = h[0], nv = h[1];
    if (G == 0) return;
    DBuf<int> nitems(G), ibase(G), bad(G);
    DBuf<unsigned long long> taken(1);
    BTK_HIP(hipMemsetAsync(taken.p, 0, sizeof(unsigned long long), s));
    BTK_HIP(hipMemsetAsync(bad.p, 0, sizeof(int) * G, s));
    k_cc_verify<<<nblk(nv), TB, 0, s>>>(nv, vals2.p, goff_g, pct, meta, gid.p, gstart.p, level_boxes(tt),
                                        level_boxes(ts), bad.p);
    BTK_LAUNCH_CHECK();
    k_cc_count<<<nblk(G), TB, 0, s>>>(G, gstart.p, nvalid.p, bad.p, min_group, nitems.p, taken.p);
    BTK_LAUNCH_CHECK();
    exclusive_sum_i32(nitems.p, ibase.p, G, s);
    long long *hl = host_scratch().v + 2;
    BTK_HIP(hipMemcpyAsync(h, ibase.p + (G - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipMemcpyAsync(h + 1, nitems.p + (G - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipMemcpyAsync(hl, taken.p, sizeof(long long), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    const int total = h[0] + h[1];
    if (total == 0) return;
    cg.shared = *hl;
    cg.items.resize(total);
    cg.ent_T.resize(nv);
    cg.ent_S.resize(nv);
    k_cc_fill<<<G, TB, 0, s>>>(G, gstart.p, nvalid.p, nitems.p, ibase.p, cg.items.p);
    BTK_LAUNCH_CHECK();
    k_cc_entries<<<nblk(nv), TB, 0, s>>>(nv, vals2.p, goff_g, pct, gid.p, nitems.p, meta, cg.ent_T.p, cg.ent_S.p);
    BTK_LAUNCH_CHECK();
    {  // the covered entries of each target panel, for k_cc_reduce
        DBuf<int> tk(nv), tk2(nv), kv(nv);
        cg.tgt_k.resize(nv);
        k_cc_taken_keys<<<nblk(nv), TB, 0, s>>>(nv, gid.p, nitems.p, cg.ent_T.p, pct, tk.p, kv.p);
        BTK_LAUNCH_CHECK();
        int pbits = 1;
        while ((1ll << pbits) <= pct) pbits++;
        sort_pairs_i32(tk.p, tk2.p, kv.p, cg.tgt_k.p, nv, pbits, s);
        cg.tgt_off.resize((size_t)pct + 1);
        k_csr_offsets<<<nblk(pct + 1), TB, 0, s>>>(nv, tk2.p, pct, cg.tgt_off.p);
        BTK_LAUNCH_CHECK();
    }
    cg.nv = nv;
    {
        DBuf<int> ik(total), ik2(total), iv(total), iv2(total);
        DBuf<int4> unsorted(total);
        k_cc_item_keys<<<nblk(total), TB, 0, s>>>(total, cg.items.p, cg.ent_T.p, ik.p, iv.p);
        BTK_LAUNCH_CHECK();
        int tbits = 1;
        while ((1ll << tbits) < pct) tbits++;
        sort_pairs_i32(ik.p, ik2.p, iv.p, iv2.p, total, tbits, s);
        BTK_HIP(hipMemcpyAsync(unsorted.p, cg.items.p, sizeof(int4) * total, hipMemcpyDeviceToDevice, s));
        k_cc_item_gather<<<nblk(total), TB, 0, s>>>(total, iv2.p, unsorted.p, cg.items.p);
        BTK_LAUNCH_CHECK();
    }
    cg.n_items = total;
}

void weights_to_f16(int pc, int P, int Ps, const double *W, unsigned short *Wh, unsigned short *Wl, int *wexp,
                    hipStream_t s) {
    if (pc == 0) return;
    BTK_REQUIRE(Ps % 64 == 0 && Ps >= P, "split weight rows: a multiple of 64 at least P long");
    k_weights_f16<<<(pc + TB / 64 - 1) / (TB / 64), TB, 0, s>>>(pc, P, Ps, W, Wh, Wl, wexp);
    BTK_LAUNCH_CHECK();
}

void launch_ccgemm(const CCGemm &cg, const Tree &tt, const Tree &ts, const unsigned short *Wh,
                   const unsigned short *Wl, const int *wexp, int n1, const NodeTables &tabs, const KParams &kp,
                   int kernel, float *stage, double *Psi, size_t comp_stride, hipStream_t s) {
    if (cg.n_items == 0) return;
    const int P = n1 * n1 * n1;
    const CcWeights wq{Wh, Wl, wexp, cc_row_stride(P)};
    const int Pst = cc_stage_stride(P);
    if (kernel == BTK_KERNEL_POT)
        launch_cc_gemm<BTK_KERNEL_POT>(cg, tt, ts, wq, n1, tabs, kp, stage, Pst, s);
    else if (kernel == BTK_KERNEL_POT_GRAD)
        launch_cc_gemm<BTK_KERNEL_POT_GRAD>(cg, tt, ts, wq, n1, tabs, kp, stage, Pst, s);
    else
        throw Error("shared-matrix CC products: potential or potential+gradient kernels");
    const int C = kernel == BTK_KERNEL_POT ? 1 : 4;
    k_cc_reduce<<<tt.pc, TB, 0, s>>>(C, P, Pst, cg.tgt_off.p, cg.tgt_k.p, stage, Psi, comp_stride);
    BTK_LAUNCH_CHECK();
}

}  // namespace btk
