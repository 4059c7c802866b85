// Work-item construction and dispatch of the interaction kernels
// (btk_interact.hpp), and the FP64 direct sum (direct_sum_impl.hpp:9-94).
#include "btk_interact.hpp"

namespace btk {
namespace {

constexpr int TB = 256;
constexpr int kMaxSplit = 16;
// Default work-item targets: 64 (near) and 32 (far) wavefronts per SIMD of
// 256 CUs; finer items balance the uneven leaf and cluster pairs.
constexpr long long kNearItems = 65536;
constexpr long long kFarItems = 32768;
inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }

// Work items per target panel: blocks of `per` targets times entry slices.
// With meta, entries whose source count is zero (taken by another kernel)
// are not counted, and a panel left without entries gets no work item.
__global__ void k_task_count(int pct, const int *__restrict__ goff, int grp, int far, const int *__restrict__ tcount,
                             int P, int per, int nsplit, const int4 *__restrict__ meta, int *__restrict__ ntask) {
    const int p = blockIdx.x * TB + threadIdx.x;
    if (p >= pct) return;
    const int g0 = goff[grp * pct + p], g1 = goff[grp * pct + p + 1];
    int ne = g1 - g0;
    if (meta) {
        int nz = 0;
        for (int e = g0; e < g1; e++) nz += meta[e].z > 0;
        ne = nz;
    }
    const int nt = far ? P : tcount[p];
    const int nb = (ne > 0 && nt > 0) ? (nt + per - 1) / per : 0;
    ntask[p] = nb * min(nsplit, ne);
}

// Source descriptor of every grouped entry (EntryView); a cluster source has
// Pd[degree index] proxies.
__global__ void k_entry_meta(int n, const int *__restrict__ gsrc, const unsigned char *__restrict__ gtype,
                             const int *__restrict__ start, const int *__restrict__ count, int4 Pd,
                             int4 *__restrict__ meta) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int S = gsrc[i];
    const int ty = gtype[i], di = ty >> 2;
    const int P = di == 0 ? Pd.x : (di == 1 ? Pd.y : (di == 2 ? Pd.z : Pd.w));
    meta[i] = (ty & 1) ? make_int4(S, 0, P, ty) : make_int4(S, start[S], count[S], ty);
}

// Mixed precision: flags the near entries (the first ndeg groups) whose boxes
// are separated enough for relative FP32 coordinates.  The target panel of
// entry e is found in goff (ascending).
template <int DIM>
__global__ void k_entry_rel(int n, int pct, int ndeg, const int *__restrict__ goff, TgtView tv, SrcView sv,
                            int4 *__restrict__ meta) {
    const int e = blockIdx.x * TB + threadIdx.x;
    const int nk = ndeg * pct;
    if (e >= n || e >= goff[nk]) return;
    int lo = 0, hi = nk;  // goff[lo] <= e < goff[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (goff[mid] <= e) lo = mid;
        else hi = mid;
    }
    const int T = lo % pct;
    if (separated_for_rel<DIM>(T, meta[e].x, tv, sv)) meta[e].w |= kRelEntry;
}

// Entry slices balanced by source count (a leaf's points, or P proxies).
// One wavefront per panel: the entries' source counts and their running sum
// go to LDS, lane 0 places the slice boundaries, the lanes write the items.
// Slice k closes after the first entry at which the running count reaches
// (k + 1) / ns of the total, provided each later slice keeps at least one
// entry; when no entry qualifies, slice k runs to the end and the later
// slices are empty.
constexpr int kTaskEntries = 4096;  // entries of one panel held in LDS

__device__ __forceinline__ long long wave_incl_scan(long long v) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int d = 1; d < kLanes; d <<= 1) {
        const long long u = __shfl_up(v, d, kLanes);
        if (lane >= d) v += u;
    }
    return v;
}

__global__ void __launch_bounds__(kLanes)
    k_task_fill(int pct, const int *__restrict__ goff, int grp, int far, const int *__restrict__ gsrc,
                const unsigned char *__restrict__ gtype, const int *__restrict__ scount, int P, int per, int nsplit,
                const int4 *__restrict__ meta, const int *__restrict__ tcount, const int *__restrict__ ntask,
                const int *__restrict__ tbase, int *__restrict__ tpanel, int *__restrict__ tfirst,
                int *__restrict__ te0, int *__restrict__ te1) {
    __shared__ long long cum[kTaskEntries];
    __shared__ int bound[kMaxSplit + 1];
    __shared__ int s_ns;
    const int p = blockIdx.x;
    const int lane = threadIdx.x;
    if (ntask[p] == 0) return;
    const int g0 = goff[grp * pct + p], g1 = goff[grp * pct + p + 1];
    const int ne_all = g1 - g0;
    auto nsrc = [&](int e) -> long long { return meta ? meta[e].z : ((gtype[e] & 1) ? P : scount[gsrc[e]]); };
    long long carry = 0;
    int nz = 0;
    for (int c0 = 0; c0 < ne_all; c0 += kLanes) {
        const int i = c0 + lane;
        const long long v = i < ne_all ? nsrc(g0 + i) : 0;
        const long long sc = wave_incl_scan(v);
        if (i < ne_all && i < kTaskEntries) cum[i] = carry + sc;
        carry += __shfl(sc, kLanes - 1, kLanes);
        nz += __popcll(__ballot(i < ne_all && (!meta || v > 0)));
    }
    const long long total = carry;
    if (lane == 0) {
        const int ns = min(nsplit, nz);
        bound[0] = g0;
        int k = 0, prev = -1;  // prev: index of the entry that closed the last slice
        for (; k < ns - 1; k++) {
            int i = -1;
            if (ne_all <= kTaskEntries) {
                int lo = prev + 1, hi = ne_all;  // first i >= lo with cum[i] * ns >= (k + 1) * total
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (cum[mid] * ns >= (k + 1) * total) hi = mid;
                    else lo = mid + 1;
                }
                i = lo < ne_all ? lo : -1;
            } else {  // more entries than LDS holds: running sum from memory
                long long c = 0;
                for (int j = 0; j < ne_all; j++) {
                    c += nsrc(g0 + j);
                    if (j > prev && c * ns >= (k + 1) * total) {
                        i = j;
                        break;
                    }
                }
            }
            if (i < 0 || (g1 - (g0 + i + 1)) < (ns - 1 - k)) break;
            bound[k + 1] = g0 + i + 1;
            prev = i;
        }
        for (k = k + 1; k <= ns; k++) bound[k] = g1;  // the open slice runs to g1, later ones are empty
        s_ns = ns;
    }
    __syncthreads();
    const int ns = s_ns;
    const int nt = far ? P : tcount[p];
    const int nb = (nt + per - 1) / per;
    const int base = tbase[p];
    for (int i = lane; i < nb * ns; i += kLanes) {
        const int bb = i / ns, k = i % ns;
        tpanel[base + i] = p;
        tfirst[base + i] = bb * per;
        te0[base + i] = bound[k];
        te1[base + i] = bound[k + 1];
    }
}

// Estimated cost of every work item of a queue (targets x sources of its
// entry slice), as a key that sorts the biggest first, and its flattened
// number as the value.
__global__ void k_queue_cost(int n, int ngroups, const QueueGroup *__restrict__ groups, int dim,
                             const int4 *__restrict__ meta, const int *__restrict__ tcount,
                             const int *__restrict__ tpt, unsigned long long *__restrict__ key,
                             int *__restrict__ val) {
    const int f = blockIdx.x * TB + threadIdx.x;
    if (f >= n) return;
    int g = 0;
    while (g + 1 < ngroups && groups[g + 1].first <= f) g++;
    const QueueGroup &G = groups[g];
    const int k = f - G.first;
    const int P = dim == 3 ? G.n1 * G.n1 * G.n1 : G.n1 * G.n1;
    const int nt = G.far ? P : tcount[G.tk.panel[k]];
    const int per = kLanes * tpt[G.far];
    const long long targets = min(per, nt - G.tk.first[k]);
    long long sources = 0;
    for (int e = G.tk.e0[k]; e < G.tk.e1[k]; e++) sources += meta[e].z;
    key[f] = ~(unsigned long long)(targets * sources);
    val[f] = f;
}

__global__ void k_queue_order(int n, int ngroups, const QueueGroup *__restrict__ groups, const int *__restrict__ val,
                              int2 *__restrict__ order) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int f = val[i];
    int g = 0;
    while (g + 1 < ngroups && groups[g + 1].first <= f) g++;
    order[i] = make_int2(g, f - groups[g].first);
}

template <int DIM, int KERNEL>
void launch_all(const InteractPlan &p, const double *q_sorted, const double *W, double *out, double *Psi,
                hipStream_t s, KernelEvents *kev) {
    const Tree &tt = *p.tt, &ts = *p.ts;
    if (kev) kev->cc_shared = p.cc_shared;
    const bool queued = p.queue.n > 0;
    if (queued) {
        // every PP/PC/CP/CC work item in one persistent launch ("near" time);
        // the shared-matrix CC products and mutual pairs follow ("far" time)
        QueueLaunch Q;
        Q.q = &p.queue;
        Q.ev = EntryView{p.meta.p};
        Q.tv = TgtView{tt.start.p, tt.count.p, p.sorted_out ? nullptr : tt.perm.p, tt.sx.p, tt.sy.p, tt.sz.p,
                       tt.lo[0].p, tt.lo[1].p, tt.lo[2].p, tt.hi[0].p, tt.hi[1].p, tt.hi[2].p};
        Q.sv = SrcView{ts.start.p, ts.count.p, ts.sx.p, ts.sy.p, ts.sz.p, q_sorted, ts.lo[0].p, ts.lo[1].p,
                       ts.lo[2].p, ts.hi[0].p, ts.hi[1].p, ts.hi[2].p, W};
        Q.kp = p.kp;
        Q.same_tree = &tt == &ts ? 1 : 0;
        Q.excl = p.o.exclude_self ? 1 : 0;
        Q.f32 = p.o.precision == BTK_PREC_MIXED ? 1 : 0;
        Q.W = W;
        Q.out = out;
        Q.Psi = Psi;
        if (kev) BTK_HIP(hipEventRecord(kev->ev[0], s));
        launch_queue<DIM, KERNEL>(Q, s);
        if (kev) {
            BTK_HIP(hipEventRecord(kev->ev[1], s));
            kev->near_done = true;
            BTK_HIP(hipEventRecord(kev->ev[2], s));
        }
    }
    for (int di = 0; di < p.ndeg; di++)  // the CC products' A operand
        if (p.g[1][di].cc.n_items) {
            const int P = ipow(p.deg[di] + 1, DIM);
            weights_to_f16(p.ts->pc, P, cc_row_stride(P), W + p.woff[di], p.wh.p + p.w16off[di], p.wl.p + p.w16off[di],
                           p.wexp.p + (size_t)di * p.ts->pc, s);
        }
    for (int far = 0; far < 2; far++) {
        bool any = false;
        for (int di = 0; di < p.ndeg; di++) any |= p.g[far][di].present;
        if (!any) continue;
        if (kev && !queued) BTK_HIP(hipEventRecord(kev->ev[far ? 2 : 0], s));
        for (int di = 0; di < p.ndeg; di++) {
            const InteractPlan::Group &g = p.g[far][di];
            if (!g.present) continue;
            const int n1 = p.deg[di] + 1;
            const int P = ipow(n1, DIM);
            double *dst = far ? Psi + p.poff[di] : out;
            InteractLaunch L;
            L.tk = TaskView{g.tk.panel.p, g.tk.first.p, g.tk.e0.p, g.tk.e1.p};
            L.ev = EntryView{p.meta.p};
            L.tv = TgtView{tt.start.p, tt.count.p, p.sorted_out ? nullptr : tt.perm.p, tt.sx.p, tt.sy.p, tt.sz.p,
                           tt.lo[0].p, tt.lo[1].p, tt.lo[2].p, tt.hi[0].p, tt.hi[1].p, tt.hi[2].p};
            L.sv = SrcView{ts.start.p, ts.count.p, ts.sx.p, ts.sy.p, ts.sz.p, q_sorted, ts.lo[0].p, ts.lo[1].p,
                           ts.lo[2].p, ts.hi[0].p, ts.hi[1].p, ts.hi[2].p, W + p.woff[di]};
            L.n_tasks = g.tk.n;
            L.n1 = n1;
            L.tpt = g.tpt;
            L.waves = g.waves;
            L.kahan = p.o.summation == BTK_SUM_KAHAN ? 1 : 0;
            L.far = far;
            L.excl = p.o.exclude_self ? 1 : 0;
            L.f32 = p.o.precision == BTK_PREC_MIXED ? 1 : 0;
            L.atomic = g.nsplit > 1 ? 1 : 0;
            // far targets are proxy points, which no zero-distance rule ties to a panel
            L.same_tree = (&tt == &ts && !far) ? 1 : 0;
            L.tabs = p.tabs[di];
            L.kp = p.kp;
            L.out = dst;
            L.comp_stride = far ? (size_t)tt.pc * P : (size_t)tt.n;
            if (g.mutual) launch_mutual<DIM, KERNEL>(g.mu, tt, q_sorted, p.kp, dst, p.sorted_out, s);
            if (g.cc.n_items)
                launch_ccgemm(g.cc, tt, ts, p.wh.p + p.w16off[di], p.wl.p + p.w16off[di],
                              p.wexp.p + (size_t)di * ts.pc, n1, p.tabs[di], p.kp, KERNEL, p.ccstage.p, dst,
                              L.comp_stride, s);
            if (g.tk.n && !queued) launch_interact<DIM, KERNEL>(L, s);
        }
        if (kev && !queued) {
            BTK_HIP(hipEventRecord(kev->ev[far ? 3 : 1], s));
            (far ? kev->far_done : kev->near_done) = true;
        }
    }
    if (kev && queued) {
        BTK_HIP(hipEventRecord(kev->ev[3], s));
        kev->far_done = true;
    }
}

// ---------------------------------------------------------------------------
// FP64 direct sum: one thread per target, sources streamed through LDS.
// ---------------------------------------------------------------------------
template <int DIM, int KERNEL, bool EXCL>
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
            eval_f64<DIM, KERNEL, EXCL>(tx - tile[0][k], ty - tile[1][k], DIM == 3 ? tz - tile[2][k] : 0.0,
                                        tile[3][k], kp.eps2, acc);
        __syncthreads();
    }
    if (valid) {
#pragma unroll
        for (int c = 0; c < C; c++) out[c * (size_t)m + a] += comp_scale<DIM, KERNEL>(c, kp.gfc) * acc[c];
    }
}

template <int DIM, int KERNEL>
void launch_direct(const KParams &kp, int m, const int *idx, const double *xt, const double *yt, const double *zt,
                   int n, const double *xs, const double *ys, const double *zs, const double *q, double *out,
                   hipStream_t s) {
    const int g = nblk(m);
    if (kp.excl)
        k_direct<DIM, KERNEL, true><<<g, TB, 0, s>>>(m, idx, xt, yt, zt, n, xs, ys, zs, q, kp, out);
    else
        k_direct<DIM, KERNEL, false><<<g, TB, 0, s>>>(m, idx, xt, yt, zt, n, xs, ys, zs, q, kp, out);
    BTK_LAUNCH_CHECK();
}

}  // namespace

// Work items per panel of group grp (nt) and their exclusive sum (tb);
// returns the total.
int count_tasks(const List &l, const Tree &tt, int grp, int P, int per, int nsplit, const int4 *meta, DBuf<int> &nt,
                DBuf<int> &tb, hipStream_t s) {
    const int pct = tt.pc;
    const int far = grp >= l.ndeg ? 1 : 0;
    k_task_count<<<nblk(pct), TB, 0, s>>>(pct, l.goff.p, grp, far, tt.count.p, P, per, nsplit, meta, nt.p);
    BTK_LAUNCH_CHECK();
    exclusive_sum_i32(nt.p, tb.p, pct, s);
    int *h = reinterpret_cast<int *>(host_scratch().v);
    BTK_HIP(hipMemcpyAsync(h, tb.p + (pct - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipMemcpyAsync(h + 1, nt.p + (pct - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    return h[0] + h[1];
}

void build_tasks(const List &l, const Tree &tt, const Tree &ts, int grp, int P, int per, int nsplit, Tasks &tk,
                 hipStream_t s, const int4 *meta) {
    const int pct = tt.pc;
    const int far = grp >= l.ndeg ? 1 : 0;
    DBuf<int> nt(pct), tb(pct);
    tk.n = count_tasks(l, tt, grp, P, per, nsplit, meta, nt, tb, s);
    const size_t m = tk.n ? tk.n : 1;
    tk.panel.resize(m);
    tk.first.resize(m);
    tk.e0.resize(m);
    tk.e1.resize(m);
    k_task_fill<<<pct, kLanes, 0, s>>>(pct, l.goff.p, grp, far, l.gsrc.p, l.gtype.p, ts.count.p, P, per, nsplit, meta,
                                       tt.count.p, nt.p, tb.p, tk.panel.p, tk.first.p, tk.e0.p, tk.e1.p);
    BTK_LAUNCH_CHECK();
}

btk_tuning &tuning() {
    static btk_tuning t{};
    return t;
}

// Resident workgroups of a kernel on the whole GPU (a persistent grid).
int queue_grid(const void *kernel) {
    int dev = 0, cus = 0, nb = 0;
    BTK_HIP(hipGetDevice(&dev));
    BTK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BTK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kLanes, 0));
    return std::max(1, nb) * std::max(1, cus);
}

int choose_split(long long base_items, long long target_items) {
    if (base_items <= 0) return 1;
    long long k = (target_items + base_items - 1) / base_items;
    return (int)std::max(1ll, std::min<long long>(kMaxSplit, k));
}

// The queue of BTK_SCHEDULE_QUEUE: the work items of every group, biggest
// first.  Near groups of several degrees write the same outputs, so their
// items then add atomically.
void build_queue(InteractPlan &p, const List &l, const Tree &tt, hipStream_t s) {
    BTK_REQUIRE(p.o.schedule == BTK_SCHEDULE_QUEUE, "not a queued plan");
    std::vector<QueueGroup> hg;
    int total = 0, near_groups = 0;
    for (int di = 0; di < p.ndeg; di++) near_groups += p.g[0][di].tk.n > 0;
    for (int far = 0; far < 2; far++)
        for (int di = 0; di < p.ndeg; di++) {
            const InteractPlan::Group &g = p.g[far][di];
            if (g.tk.n == 0) continue;
            const int n1 = p.deg[di] + 1;
            QueueGroup q{};
            q.tk = TaskView{g.tk.panel.p, g.tk.first.p, g.tk.e0.p, g.tk.e1.p};
            q.woff = p.woff[di];
            q.ooff = far ? p.poff[di] : 0;
            q.comp_stride = far ? (size_t)tt.pc * ipow(n1, tt.dim) : (size_t)tt.n;
            q.n1 = n1;
            q.far = far;
            q.atomic = (g.nsplit > 1 || (!far && near_groups > 1)) ? 1 : 0;
            q.first = total;
            q.tabs = p.tabs[di];
            (far ? p.queue.ftpt : p.queue.ntpt) = g.tpt;
            total += g.tk.n;
            hg.push_back(q);
        }
    p.queue.n = total;
    if (total == 0) return;
    const int ng = (int)hg.size();
    p.queue.groups.resize(ng);
    p.queue.order.resize(total);
    p.queue.ticket.resize(1);
    BTK_HIP(hipMemcpyAsync(p.queue.groups.p, hg.data(), sizeof(QueueGroup) * ng, hipMemcpyHostToDevice, s));
    const int htpt[2] = {p.queue.ntpt, p.queue.ftpt};
    DBuf<int> tpt(2);
    BTK_HIP(hipMemcpyAsync(tpt.p, htpt, sizeof htpt, hipMemcpyHostToDevice, s));
    DBuf<unsigned long long> key(total), key2(total);
    DBuf<int> val(total), val2(total);
    k_queue_cost<<<nblk(total), TB, 0, s>>>(total, ng, p.queue.groups.p, tt.dim, p.meta.p, tt.count.p, tpt.p, key.p,
                                            val.p);
    BTK_LAUNCH_CHECK();
    sort_pairs_u64(key.p, key2.p, val.p, val2.p, total, 64, s);
    k_queue_order<<<nblk(total), TB, 0, s>>>(total, ng, p.queue.groups.p, val2.p, p.queue.order.p);
    BTK_LAUNCH_CHECK();
    BTK_HIP(hipStreamSynchronize(s));  // hg and htpt are host memory of this call
}

std::shared_ptr<InteractPlan> plan_interactions(const btk_run_config &cfg, const btk_options &o, const Tree &tt,
                                                const Tree &ts, const List &l, hipStream_t s, bool sorted_out) {
    BTK_REQUIRE(tt.dim == ts.dim, "target and source trees differ in dimension");
    const int dim = tt.dim;
    const int n1 = cfg.interp_degree + 1;
    BTK_REQUIRE(n1 >= 2 && n1 <= kMaxN1, "interp_degree must be between 1 and 10");
    BTK_REQUIRE(o.precision == BTK_PREC_FP64 || o.precision == BTK_PREC_MIXED, "unknown precision");
    BTK_REQUIRE(o.schedule == BTK_SCHEDULE_GROUPS || o.schedule == BTK_SCHEDULE_QUEUE, "unknown schedule");
    BTK_REQUIRE(o.summation == BTK_SUM_BLOCKED || o.summation == BTK_SUM_KAHAN, "unknown near-field summation");
    BTK_REQUIRE(!(o.summation == BTK_SUM_KAHAN && o.schedule == BTK_SCHEDULE_QUEUE),
                "Kahan-compensated near sums run with the group launches (schedule 0)");
    BTK_REQUIRE(!(dim == 2 && o.kernel == BTK_KERNEL_REF_VEL3D), "the reference vel kernel is 3D only");
    BTK_REQUIRE(l.ndeg >= 1 && l.ndeg <= 4 && l.deg[l.ndeg - 1] == cfg.interp_degree,
                "the list was built for another interp_degree");
    auto p = std::make_shared<InteractPlan>();
    p->cfg = cfg;
    p->o = o;
    p->sorted_out = sorted_out;
    p->tt = &tt;
    p->ts = &ts;
    p->kp = make_kparams(dim, cfg.ker_eps, o);
    p->ndeg = l.ndeg;
    const int C = components(dim, o.kernel);
    int Pd[4] = {0, 0, 0, 0};
    for (int di = 0; di < l.ndeg; di++) {
        p->deg[di] = l.deg[di];
        p->tabs[di] = node_tables(l.deg[di]);
        p->woff[di] = proxy_offset(l, ts.pc, 1, di);
        p->poff[di] = proxy_offset(l, tt.pc, C, di);
        Pd[di] = ipow(l.deg[di] + 1, dim);
    }
    const bool f32 = o.precision == BTK_PREC_MIXED;
    const btk_tuning &tu = tuning();
    p->meta.resize(l.count ? l.count : 1);
    if (l.count) {
        k_entry_meta<<<nblk(l.count), TB, 0, s>>>(l.count, l.gsrc.p, l.gtype.p, ts.start.p, ts.count.p,
                                                  make_int4(Pd[0], Pd[1], Pd[2], Pd[3]), p->meta.p);
        BTK_LAUNCH_CHECK();
        if (o.precision == BTK_PREC_MIXED && tt.pc > 0) {
            const TgtView tv{tt.start.p, tt.count.p, nullptr, tt.sx.p, tt.sy.p, tt.sz.p, tt.lo[0].p, tt.lo[1].p,
                             tt.lo[2].p, tt.hi[0].p, tt.hi[1].p, tt.hi[2].p};
            const SrcView sv{ts.start.p, ts.count.p, ts.sx.p, ts.sy.p, ts.sz.p, nullptr, ts.lo[0].p, ts.lo[1].p,
                             ts.lo[2].p, ts.hi[0].p, ts.hi[1].p, ts.hi[2].p, nullptr};
            if (dim == 3)
                k_entry_rel<3><<<nblk(l.count), TB, 0, s>>>(l.count, tt.pc, l.ndeg, l.goff.p, tv, sv, p->meta.p);
            else
                k_entry_rel<2><<<nblk(l.count), TB, 0, s>>>(l.count, tt.pc, l.ndeg, l.goff.p, tv, sv, p->meta.p);
            BTK_LAUNCH_CHECK();
        }
    }
    for (int far = 0; far < 2; far++) {
        if (l.counts[2 * far] + l.counts[2 * far + 1] == 0) continue;
        for (int di = 0; di < l.ndeg; di++) {
            InteractPlan::Group &g = p->g[far][di];
            const int grp = far * l.ndeg + di;
            const int P = Pd[di];
            const int tpt_set = far ? tu.far_tpt : tu.near_tpt;
            // default: two targets per lane where the kernel is short (2D), one in 3D
            g.tpt = !f32 ? 1 : (tpt_set ? tpt_set : ((dim == 3 || (far && P <= kLanes)) ? 1 : 2));
            if (!far && o.summation == BTK_SUM_KAHAN) g.tpt = 1;  // the compensated kernel: one target per lane
            // default: the 3D near launch shares each source tile between 4
            // wavefronts (a leaf holds hundreds of targets; 16M Plummer
            // points: 3.5x fewer bytes fetched at the same time); in 2D the
            // shorter kernel loses to the barriers (1M points: 9.4 against
            // 8.6 ms), and the far launch's targets are the proxies of one
            // cluster: one wavefront each
            const int waves_set = far ? tu.far_waves : tu.near_waves;
            g.waves = waves_set ? waves_set : ((far || dim == 2) ? 1 : 4);
            if (o.schedule == BTK_SCHEDULE_QUEUE) g.waves = 1;  // the persistent kernel's workgroup
            const int per = kLanes * g.tpt * g.waves;
            // mixed precision, one tree, when asked for: mutual PP pairs once
            // for both directions (PP entries are in the degree-0 near group)
            g.mutual = !far && di == 0 && f32 && &tt == &ts && tu.mutual && o.summation != BTK_SUM_KAHAN;
            if (g.mutual) build_mutual(l, tt, g.tpt, p->meta.p, g.mu, s);
            // mixed precision, 3D: CC pairs that share a kernel matrix run as
            // MFMA products
            if (far && f32 && dim == 3 && o.kernel != BTK_KERNEL_REF_VEL3D && tu.cc_group >= 0) {
                build_ccgemm(l, tt, ts, grp, tu.cc_group > 0 ? tu.cc_group : kCcMinGroup, p->meta.p, g.cc, s);
                p->cc_shared += g.cc.shared;
            }
            const int4 *meta = (g.mutual || g.cc.n_items > 0) ? p->meta.p : nullptr;
            long long base;
            if (l.ndeg == 1) {  // one degree: work items of the whole class, bounded from the tree
                const int n_clusters = tt.pc - tt.n_leaves;
                base = far ? (long long)n_clusters * ((P + per - 1) / per) : (long long)tt.n / per + tt.n_leaves;
            } else {  // several: this group's own items at one slice per panel
                DBuf<int> nt(tt.pc), tb(tt.pc);
                base = count_tasks(l, tt, grp, P, per, 1, meta, nt, tb, s);
            }
            const long long items = far ? (tu.far_items ? tu.far_items : kFarItems)
                                        : (tu.near_items ? tu.near_items : kNearItems);
            g.nsplit = choose_split(base, items);
            build_tasks(l, tt, ts, grp, P, per, g.nsplit, g.tk, s, meta);
            g.present = g.tk.n > 0 || g.mutual || g.cc.n_items > 0;
        }
    }
    size_t n16 = 0, nstage = 0;
    for (int di = 0; di < l.ndeg; di++) {
        p->w16off[di] = n16;
        const CCGemm &cc = p->g[1][di].cc;
        if (cc.n_items > 0) {
            const int P = ipow(l.deg[di] + 1, dim);
            n16 += (size_t)ts.pc * cc_row_stride(P);
            nstage = std::max(nstage, (size_t)cc.nv * C * cc_stage_stride(P));
        }
    }
    if (nstage) p->ccstage.resize(nstage);
    if (n16) {
        p->wh.resize(n16);
        p->wl.resize(n16);
        p->wexp.resize((size_t)ts.pc * l.ndeg);
    }
    if (o.schedule == BTK_SCHEDULE_QUEUE) build_queue(*p, l, tt, s);
    return p;
}

void launch_interactions(const InteractPlan &p, const double *q_sorted, const double *W, double *out, double *Psi,
                         hipStream_t s, KernelEvents *kev) {
    const int dim = p.tt->dim, kernel = p.o.kernel;
    if (dim == 3) {
        if (kernel == BTK_KERNEL_POT)
            launch_all<3, BTK_KERNEL_POT>(p, q_sorted, W, out, Psi, s, kev);
        else if (kernel == BTK_KERNEL_REF_VEL3D)
            launch_all<3, BTK_KERNEL_REF_VEL3D>(p, q_sorted, W, out, Psi, s, kev);
        else
            launch_all<3, BTK_KERNEL_POT_GRAD>(p, q_sorted, W, out, Psi, s, kev);
    } else {
        if (kernel == BTK_KERNEL_POT)
            launch_all<2, BTK_KERNEL_POT>(p, q_sorted, W, out, Psi, s, kev);
        else
            launch_all<2, BTK_KERNEL_POT_GRAD>(p, q_sorted, W, out, Psi, s, kev);
    }
}

void fmm_interactions(const btk_run_config &cfg, const btk_options &o, const Tree &tt, const Tree &ts, const List &l,
                      const double *q_sorted, const double *W, double *out, double *Psi, hipStream_t s,
                      KernelEvents *kev, bool sorted_out) {
    // the plan's buffers go back to the cache on return, while the launches
    // may still run; the cache hands them out only to this thread's later
    // work on this stream until the API call has synchronized
    const auto p = plan_interactions(cfg, o, tt, ts, l, s, sorted_out);
    launch_interactions(*p, q_sorted, W, out, Psi, s, kev);
}

void direct_sum(int dim, const btk_options &o, double eps, int m, const int *idx, const double *xt, const double *yt,
                const double *zt, int n, const double *xs, const double *ys, const double *zs, const double *q,
                double *out, hipStream_t s) {
    if (m <= 0) return;
    const KParams kp = make_kparams(dim, eps, o);
    if (dim == 3) {
        if (o.kernel == BTK_KERNEL_POT)
            launch_direct<3, BTK_KERNEL_POT>(kp, m, idx, xt, yt, zt, n, xs, ys, zs, q, out, s);
        else if (o.kernel == BTK_KERNEL_REF_VEL3D)
            launch_direct<3, BTK_KERNEL_REF_VEL3D>(kp, m, idx, xt, yt, zt, n, xs, ys, zs, q, out, s);
        else
            launch_direct<3, BTK_KERNEL_POT_GRAD>(kp, m, idx, xt, yt, zt, n, xs, ys, zs, q, out, s);
    } else {
        BTK_REQUIRE(o.kernel != BTK_KERNEL_REF_VEL3D, "the reference vel kernel is 3D only");
        if (o.kernel == BTK_KERNEL_POT)
            launch_direct<2, BTK_KERNEL_POT>(kp, m, idx, xt, yt, nullptr, n, xs, ys, nullptr, q, out, s);
        else
            launch_direct<2, BTK_KERNEL_POT_GRAD>(kp, m, idx, xt, yt, nullptr, n, xs, ys, nullptr, q, out, s);
    }
}

}  // namespace btk
