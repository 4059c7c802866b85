// Dual-tree traversal on the GPU (the reference's dual_tree_traversal_2d/3d,
// interaction_list.cpp:5-250, is a host-serial FIFO walk).
//
// The FIFO walk visits pairs generation by generation: the pairs pushed while
// generation k is processed are exactly generation k+1, in order.  So each
// generation is one array; a classify kernel applies the MAC and the refine
// rule to every pair, a scan gives every emitted pair its slot in the output
// list and every refined pair the slots of its 2^dim children in the next
// generation, and the output list is the reference's list in the reference's
// order.  The MAC geometry (box centres, distance, ratio) is evaluated with
// the reference's expressions in FP64 without contraction (-ffp-contract=off).
#include <algorithm>

#include "btk_mac.hpp"

namespace btk {
namespace {

constexpr int TB = 256;

inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }

struct TreeView {
    const double *lo0, *lo1, *lo2, *hi0, *hi1, *hi2, *radius;
    const int *count, *child0;
};

TreeView view(const Tree &t) {
    return TreeView{t.lo[0].p, t.lo[1].p, t.lo[2].p, t.hi[0].p, t.hi[1].p, t.hi[2].p,
                    t.radius.p, t.count.p, t.child0.p};
}

// Classification codes: below kRefineT an emitted pair, type | degree index
// << 2 (the degree index is 0 but for the learned MAC); kRefineT refine the
// target, kRefineS refine the source.
constexpr int kRefineT = 16, kRefineS = 17;

template <int DIM, bool MAC>
__device__ __forceinline__ int classify_pair(int it, int is, const TreeView &T, const TreeView &S, double theta,
                                             int thr, const MacView &mac, const float *mw) {
    // interaction_list.cpp:141-154
    const double x1 = 0.5 * (T.lo0[it] + T.hi0[it]);
    const double y1 = 0.5 * (T.lo1[it] + T.hi1[it]);
    const double x2 = 0.5 * (S.lo0[is] + S.hi0[is]);
    const double y2 = 0.5 * (S.lo1[is] + S.hi1[is]);
    const double ex = x2 - x1, ey = y2 - y1;
    double dist;
    if constexpr (DIM == 3) {
        const double z1 = 0.5 * (T.lo2[it] + T.hi2[it]);
        const double z2 = 0.5 * (S.lo2[is] + S.hi2[is]);
        const double ez = z2 - z1;
        dist = sqrt(ex * ex + ey * ey + ez * ez);
    } else {
        dist = sqrt(ex * ex + ey * ey);
    }
    double sep = 100;
    if (dist > 0) sep = (T.radius[it] + S.radius[is]) / dist;
    const int ct = T.count[it], cs = S.count[is];
    const int type = (ct > thr ? 2 : 0) + (cs > thr ? 1 : 0);
    const bool accepted = sep < theta;  // :155-179
    if constexpr (DIM == 3 && MAC) {
        // the learned MAC picks the type and degree of an accepted pair, and
        // may accept what the theta MAC would refine
        const int c = mac_decide(it, is, mac, mw, thr, type, accepted);
        if (c >= 0) return c;
    }
    if (accepted) return type;
    // :180-203
    const bool lt = T.child0[it] < 0, lsf = S.child0[is] < 0;
    if (lt && lsf) return 0;
    if (lt) return kRefineS;
    if (lsf) return kRefineT;
    return ct > cs ? kRefineT : kRefineS;
}

template <int DIM, bool MAC>
__global__ void __launch_bounds__(TB) k_classify(int G, const int2 *__restrict__ cur, TreeView T, TreeView S,
                                                 double theta, int thr, MacView mac, unsigned char *__restrict__ code,
                                                 long long *__restrict__ val) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= G) return;
    const int c = classify_pair<DIM, MAC>(cur[i].x, cur[i].y, T, S, theta, thr, mac, mac.w);
    code[i] = (unsigned char)c;
    val[i] = c < kRefineT ? 1ll : (1ll << 32);
}

// Emitted pair or refined children of pair i, given ex = emitted count |
// refined count << 32 of the pairs before it.
template <int DIM>
__device__ __forceinline__ void write_pair(int c, int it, int is, long long ex, const TreeView &T,
                                           const TreeView &S, int lbase, btk_interact_pair *__restrict__ out,
                                           int2 *__restrict__ next) {
    constexpr int NC = DIM == 3 ? 8 : 4;
    if (c < kRefineT) {
        btk_interact_pair p;
        p.target_panel = it;
        p.source_panel = is;
        p.interact_type = c;
        out[lbase + (int)(ex & 0xffffffffLL)] = p;
    } else {
        const int k = (int)(ex >> 32);
        // :204-238: children pushed in c1..c8 order with the other side repeated
        if (c == kRefineT) {
            const int b = T.child0[it];
            for (int j = 0; j < NC; j++) next[NC * k + j] = make_int2(b + j, is);
        } else {
            const int b = S.child0[is];
            for (int j = 0; j < NC; j++) next[NC * k + j] = make_int2(it, b + j);
        }
    }
}

// A small generation in one workgroup: classify, exclusive scan of the
// packed counts, writes; the totals (emitted | refined << 32) go to *tot in
// pinned host memory.
constexpr int kFuseMax = 8192;

// Threads of the fused generation: 1024, or 512 with the learned MAC, whose
// network needs the registers of two wavefronts per SIMD.
template <bool MAC>
constexpr int fused_threads() {
    return MAC ? 512 : 1024;
}

template <int DIM, bool MAC, int FB = fused_threads<MAC>()>
__global__ void __launch_bounds__(FB)
    k_generation(int G, const int2 *__restrict__ cur, TreeView T, TreeView S, double theta, int thr, MacView mac,
                 int lbase, btk_interact_pair *__restrict__ out, int2 *__restrict__ next, long long *tot) {
    __shared__ long long wsum[FB / 64];
    __shared__ long long btot;
    __shared__ float mw[DIM == 3 && MAC ? kMacParams : 1];  // the learned MAC's weights
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    if constexpr (DIM == 3 && MAC) {
        for (int i = tid; i < kMacParams; i += FB) mw[i] = mac.w[i];
        __syncthreads();
    }
    long long carry = 0;
    for (int c0 = 0; c0 < G; c0 += FB) {
        const int i = c0 + tid;
        int c = 0, it = 0, is = 0;
        long long v = 0;
        if (i < G) {
            it = cur[i].x;
            is = cur[i].y;
            c = classify_pair<DIM, MAC>(it, is, T, S, theta, thr, mac, mw);
            v = c < kRefineT ? 1ll : (1ll << 32);
        }
        long long incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const long long u = __shfl_up(incl, d, 64);
            if (lane >= d) incl += u;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        if (w == 0) {
            long long x = lane < FB / 64 ? wsum[lane] : 0;
#pragma unroll
            for (int d = 1; d < FB / 64; d <<= 1) {
                const long long u = __shfl_up(x, d, 64);
                if (lane >= d) x += u;
            }
            if (lane < FB / 64) wsum[lane] = x;  // inclusive over wavefronts
            if (lane == FB / 64 - 1) btot = x;
        }
        __syncthreads();
        const long long ex = carry + (w > 0 ? wsum[w - 1] : 0) + incl - v;
        if (i < G) write_pair<DIM>(c, it, is, ex, T, S, lbase, out, next);
        carry += btot;
        __syncthreads();  // wsum and btot are rewritten by the next chunk
    }
    if (tid == 0) *tot = carry;
}

template <int DIM>
__global__ void __launch_bounds__(TB)
    k_write(int G, const int2 *__restrict__ cur, const unsigned char *__restrict__ code,
            const long long *__restrict__ val, const long long *__restrict__ inc, TreeView T, TreeView S, int lbase,
            btk_interact_pair *__restrict__ out, int2 *__restrict__ next) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= G) return;
    write_pair<DIM>(code[i], cur[i].x, cur[i].y, inc[i] - val[i], T, S, lbase, out, next);
}

// Keys for the grouped forms: (group << tbits) | target, group = far * ndeg +
// degree index.
// Evaluations per type use the entry's degree: Pd[degree index] proxies.
__global__ void k_group_keys(int L, const btk_interact_pair *__restrict__ pairs, int tbits, int pct, int ndeg,
                             int *__restrict__ keys, int *__restrict__ vals, int *__restrict__ cnt,
                             const int *__restrict__ tcount, const int *__restrict__ scount, int4 Pd,
                             double *__restrict__ evals) {
    const int i = blockIdx.x * TB + threadIdx.x;
    __shared__ double se[4];
    if (threadIdx.x < 4) se[threadIdx.x] = 0.0;
    __syncthreads();
    if (i < L) {
        const btk_interact_pair p = pairs[i];
        const int ty = p.interact_type & 3, di = p.interact_type >> 2;
        const int P = di == 0 ? Pd.x : (di == 1 ? Pd.y : (di == 2 ? Pd.z : Pd.w));
        const int grp = (ty >= 2 ? ndeg : 0) + di;
        keys[i] = (grp << tbits) | p.target_panel;
        vals[i] = i;
        atomicAdd(&cnt[grp * pct + p.target_panel], 1);
        const double nt = ty >= 2 ? (double)P : (double)tcount[p.target_panel];
        const double ns = (ty & 1) ? (double)P : (double)scount[p.source_panel];
        atomicAdd(&se[ty], nt * ns);
    }
    __syncthreads();
    if (threadIdx.x < 4 && se[threadIdx.x] != 0.0) atomicAdd(&evals[threadIdx.x], se[threadIdx.x]);
}

__global__ void k_group_gather(int L, const int *__restrict__ order, const btk_interact_pair *__restrict__ pairs,
                               int *__restrict__ gsrc, unsigned char *__restrict__ gtype) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= L) return;
    const btk_interact_pair p = pairs[order[i]];
    gsrc[i] = p.source_panel;
    gtype[i] = (unsigned char)p.interact_type;
}

__global__ void k_type_count(int L, const btk_interact_pair *__restrict__ pairs, unsigned long long *__restrict__ c) {
    __shared__ unsigned int sc[4];
    if (threadIdx.x < 4) sc[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i < L) atomicAdd(&sc[pairs[i].interact_type & 3], 1u);
    __syncthreads();
    if (threadIdx.x < 4 && sc[threadIdx.x]) atomicAdd(&c[threadIdx.x], (unsigned long long)sc[threadIdx.x]);
}

}  // namespace

std::unique_ptr<List> build_list(btk_run_config *cfg, const Tree &tt, const Tree &ts, hipStream_t s, bool learned_mac,
                                 std::vector<int2> *visited) {
    BTK_REQUIRE(tt.dim == ts.dim, "target and source trees differ in dimension");
    const int dim = tt.dim;
    const int NC = dim == 3 ? 8 : 4;
    auto l = std::make_unique<List>();
    l->dim = dim;
    const TreeView T = view(tt), S = view(ts);
    long long *hv = host_scratch().v;
    const MacView mac = learned_mac ? mac_view(*cfg, tt, ts) : MacView{};
    // the learned MAC picks each pair's degree among 4, 6, 8, 10 up to the
    // configured one; otherwise every pair has the configured degree
    if (learned_mac) {
        l->ndeg = mac.di + 1;
        for (int k = 0; k < l->ndeg; k++) l->deg[k] = 4 + 2 * k;
    } else {
        l->deg[0] = cfg->interp_degree;
    }

    // generation 0: the root pair of every problem (one pair for one tree)
    BTK_REQUIRE(tt.nprob == ts.nprob, "target and source forests hold different numbers of problems");
    std::vector<int2> roots(tt.nprob);
    for (int p = 0; p < tt.nprob; p++) roots[p] = make_int2(p, p);
    DBuf<int2> cur(tt.nprob), nxt;
    BTK_HIP(hipMemcpyAsync(cur.p, roots.data(), sizeof(int2) * tt.nprob, hipMemcpyHostToDevice, s));
    int G = tt.nprob;
    int L = 0;
    size_t lcap = std::max<size_t>(1024, 64 * (size_t)std::max(tt.pc, ts.pc));
    l->pairs.resize(lcap);
    auto reserve = [&](size_t need) {
        if (need <= lcap) return;
        size_t ncap = lcap;
        while (ncap < need) ncap *= 2;
        l->pairs.grow(ncap, s);
        lcap = ncap;
    };
    while (G > 0) {
        if (visited) {  // training samples: every visited pair
            const size_t v0 = visited->size();
            visited->resize(v0 + G);
            BTK_HIP(hipMemcpyAsync(visited->data() + v0, cur.p, sizeof(int2) * G, hipMemcpyDeviceToHost, s));
            BTK_HIP(hipStreamSynchronize(s));
        }
        long long tot;
        if (G <= kFuseMax) {
            // one launch; at most G pairs emitted and NC * G children
            reserve((size_t)L + G);
            nxt.resize((size_t)NC * G);
#define BTK_GEN(D, M)                                                                                      \
    k_generation<D, M><<<1, fused_threads<M>(), 0, s>>>(G, cur.p, T, S, cfg->fmm_theta, cfg->fmm_cluster_thresh, mac, L, l->pairs.p, \
                                        nxt.p, hv)
            if (dim == 3 && mac.w)
                BTK_GEN(3, true);
            else if (dim == 3)
                BTK_GEN(3, false);
            else
                BTK_GEN(2, false);
#undef BTK_GEN
            BTK_LAUNCH_CHECK();
            BTK_HIP(hipStreamSynchronize(s));
            tot = hv[0];
        } else {
            DBuf<unsigned char> code(G);
            DBuf<long long> val(G), inc(G);
#define BTK_CLS(D, M)                                                                                     \
    k_classify<D, M><<<nblk(G), TB, 0, s>>>(G, cur.p, T, S, cfg->fmm_theta, cfg->fmm_cluster_thresh, mac, code.p, \
                                            val.p)
            if (dim == 3 && mac.w)
                BTK_CLS(3, true);
            else if (dim == 3)
                BTK_CLS(3, false);
            else
                BTK_CLS(2, false);
#undef BTK_CLS
            BTK_LAUNCH_CHECK();
            inclusive_sum_i64(val.p, inc.p, G, s);
            BTK_HIP(hipMemcpyAsync(hv, inc.p + (G - 1), sizeof(long long), hipMemcpyDeviceToHost, s));
            BTK_HIP(hipStreamSynchronize(s));
            tot = hv[0];
            const int E = (int)(tot & 0xffffffffLL);
            const long long X = tot >> 32;
            BTK_REQUIRE((long long)NC * X < (1ll << 31), "traversal generation exceeds 2^31 pairs");
            reserve((size_t)L + E);
            nxt.resize(X ? (size_t)NC * X : 1);
            if (dim == 3)
                k_write<3><<<nblk(G), TB, 0, s>>>(G, cur.p, code.p, val.p, inc.p, T, S, L, l->pairs.p, nxt.p);
            else
                k_write<2><<<nblk(G), TB, 0, s>>>(G, cur.p, code.p, val.p, inc.p, T, S, L, l->pairs.p, nxt.p);
            BTK_LAUNCH_CHECK();
        }
        const int E = (int)(tot & 0xffffffffLL);
        const long long X = tot >> 32;
        BTK_REQUIRE((long long)NC * X < (1ll << 31), "traversal generation exceeds 2^31 pairs");
        L += E;
        G = (int)(NC * X);
        std::swap(cur, nxt);
    }
    l->count = L;

    // counters (interaction_list.cpp:242-249)
    DBuf<unsigned long long> tc(4);
    BTK_HIP(hipMemsetAsync(tc.p, 0, 4 * sizeof(unsigned long long), s));
    if (L) {
        k_type_count<<<nblk(L), TB, 0, s>>>(L, l->pairs.p, tc.p);
        BTK_LAUNCH_CHECK();
    }
    BTK_HIP(hipMemcpyAsync(hv, tc.p, 4 * sizeof(long long), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 4; k++) l->counts[k] = hv[k];
    cfg->fmm_interaction_count = L;
    cfg->fmm_pp_count = (int)l->counts[0];
    cfg->fmm_pc_count = (int)l->counts[1];
    cfg->fmm_cp_count = (int)l->counts[2];
    cfg->fmm_cc_count = (int)l->counts[3];

    // target-grouped forms for the interaction kernels
    const int pct = tt.pc;
    int tbits = 1;
    while ((1 << tbits) <= pct) tbits++;
    const int ng = 2 * l->ndeg;  // groups
    int gbits = 1;
    while ((1 << gbits) < ng) gbits++;
    int Pd[4] = {0, 0, 0, 0};
    for (int k = 0; k < l->ndeg; k++) Pd[k] = ipow(l->deg[k] + 1, dim);
    DBuf<int> keys(L ? L : 1), vals(L ? L : 1), keys2(L ? L : 1), order(L ? L : 1);
    DBuf<int> cnt((size_t)ng * pct + 1);
    DBuf<double> ev(4);
    BTK_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * ((size_t)ng * pct + 1), s));
    BTK_HIP(hipMemsetAsync(ev.p, 0, sizeof(double) * 4, s));
    if (L) {
        k_group_keys<<<nblk(L), TB, 0, s>>>(L, l->pairs.p, tbits, pct, l->ndeg, keys.p, vals.p, cnt.p, tt.count.p,
                                            ts.count.p, make_int4(Pd[0], Pd[1], Pd[2], Pd[3]), ev.p);
        BTK_LAUNCH_CHECK();
        sort_pairs_i32(keys.p, keys2.p, vals.p, order.p, L, tbits + gbits, s);
    }
    l->goff.resize((size_t)ng * pct + 1);
    exclusive_sum_i32(cnt.p, l->goff.p, ng * pct + 1, s);
    l->gsrc.resize(L ? L : 1);
    l->gtype.resize(L ? L : 1);
    if (L) {
        k_group_gather<<<nblk(L), TB, 0, s>>>(L, order.p, l->pairs.p, l->gsrc.p, l->gtype.p);
        BTK_LAUNCH_CHECK();
    }
    double evh[4];
    BTK_HIP(hipMemcpyAsync(evh, ev.p, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 4; k++) l->evals[k] = evh[k];
    l->pct = pct;
    return l;
}

}  // namespace btk
