// The C ABI of include/btk.h.  Every entry point converts exceptions into a
// status code and a thread-local message.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>

#include "btk_internal.hpp"

struct btk_tree {
    std::unique_ptr<btk::Tree> t;
};
struct btk_list {
    std::unique_ptr<btk::List> l;
};
struct btk_plan {
    btk_run_config cfg{};
    btk_options o{};
    int C = 1, m = 0, n = 0;
    std::unique_ptr<btk::Tree> ts, tt_own;
    const btk::Tree *tt = nullptr;
    std::unique_ptr<btk::List> l;
    std::shared_ptr<btk::InteractPlan> ip;
    btk::DBuf<double> q, W, Psi, outs;  // outs: results in the target tree's sorted order
    hipStream_t cs = nullptr;  // capture stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const double *g_q = nullptr;  // the pointers the graph was captured with
    double *g_out = nullptr;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~btk_plan() {
        drop_graph();
        if (cs) (void)hipStreamDestroy(cs);
    }
};

namespace {

thread_local std::string g_err;

template <class F>
int guarded(F &&f) {
    int rc = 0;
    try {
        f();
    } catch (const std::exception &e) {
        g_err = e.what();
        rc = 1;
    } catch (...) {
        g_err = "unknown error";
        rc = 1;
    }
    // a failed call may leave work queued that reads blocks it freed
    if (rc) (void)hipDeviceSynchronize();
    btk::end_call();
    return rc;
}

hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

btk_options opts_or_default(const btk_options *o) {
    btk_options r;
    std::memset(&r, 0, sizeof r);
    if (o) r = *o;
    return r;
}

void check_cfg(const btk_run_config *cfg, int dim) {
    BTK_REQUIRE(cfg != nullptr, "null run config");
    BTK_REQUIRE(cfg->dim == dim, "run config dim does not match the entry point");
}

// std::stoi / std::stod as read_namelist.cpp:25,31,37,39 use them: leading
// white space, then a number; trailing text is ignored; no number is an error.
int parse_int(const std::string &s) {
    const char *b = s.c_str();
    char *e = nullptr;
    errno = 0;
    long v = std::strtol(b, &e, 10);
    if (e == b) throw btk::Error("namelist: no integer in '" + s + "'");
    if (errno == ERANGE || v > INT32_MAX || v < INT32_MIN) throw btk::Error("namelist: integer out of range");
    return (int)v;
}

double parse_double(const std::string &s) {
    const char *b = s.c_str();
    char *e = nullptr;
    errno = 0;
    double v = std::strtod(b, &e);
    if (e == b) throw btk::Error("namelist: no number in '" + s + "'");
    if (errno == ERANGE && (v == HUGE_VAL || v == -HUGE_VAL)) throw btk::Error("namelist: number out of range");
    return v;
}

void copy_panels_host(const btk::Tree &t, std::vector<double> (&lo)[3], std::vector<double> (&hi)[3],
                      std::vector<double> &rad, std::vector<int> &cnt, std::vector<int> &par,
                      std::vector<int> &c0, std::vector<int> &lvl) {
    const int pc = t.pc;
    for (int d = 0; d < 3; d++) {
        lo[d].resize(pc);
        hi[d].resize(pc);
        BTK_HIP(hipMemcpy(lo[d].data(), t.lo[d].p, sizeof(double) * pc, hipMemcpyDeviceToHost));
        BTK_HIP(hipMemcpy(hi[d].data(), t.hi[d].p, sizeof(double) * pc, hipMemcpyDeviceToHost));
    }
    rad.resize(pc);
    cnt.resize(pc);
    par.resize(pc);
    c0.resize(pc);
    lvl.resize(pc);
    BTK_HIP(hipMemcpy(rad.data(), t.radius.p, sizeof(double) * pc, hipMemcpyDeviceToHost));
    BTK_HIP(hipMemcpy(cnt.data(), t.count.p, sizeof(int) * pc, hipMemcpyDeviceToHost));
    BTK_HIP(hipMemcpy(par.data(), t.parent.p, sizeof(int) * pc, hipMemcpyDeviceToHost));
    BTK_HIP(hipMemcpy(c0.data(), t.child0.p, sizeof(int) * pc, hipMemcpyDeviceToHost));
    BTK_HIP(hipMemcpy(lvl.data(), t.level.p, sizeof(int) * pc, hipMemcpyDeviceToHost));
}

}  // namespace

extern "C" {

int btk_components(int dim, int kernel) { return btk::components(dim, kernel); }

const char *btk_last_error(void) { return g_err.c_str(); }

int btk_release_cached(void) {
    return guarded([&] { btk::release_cached(); });
}

int btk_set_tuning(const btk_tuning *t) {
    return guarded([&] {
        BTK_REQUIRE(t, "null argument");
        BTK_REQUIRE(t->near_items >= 0 && t->far_items >= 0, "work item targets must be >= 0");
        BTK_REQUIRE(t->near_tpt >= 0 && t->near_tpt <= 2 && t->far_tpt >= 0 && t->far_tpt <= 2,
                    "targets per lane must be 0 (default), 1 or 2");
        BTK_REQUIRE(t->mutual == 0 || t->mutual == 1, "mutual must be 0 or 1");
        BTK_REQUIRE(t->cc_group >= -1, "cc_group must be -1 (off), 0 (default) or a minimum group size");
        auto waves_ok = [](int w) { return w == 0 || w == 1 || w == 2 || w == 4; };
        BTK_REQUIRE(waves_ok(t->near_waves) && waves_ok(t->far_waves),
                    "wavefronts per workgroup must be 0 (default), 1, 2 or 4");
        BTK_REQUIRE(t->multi_vectors == 0 || t->multi_vectors == 32 || t->multi_vectors == 64 ||
                        t->multi_vectors == 128,
                    "charge vectors per workgroup must be 0 (default), 32, 64 or 128");
        btk::tuning() = *t;
    });
}

int btk_get_tuning(btk_tuning *t) {
    return guarded([&] {
        BTK_REQUIRE(t, "null argument");
        *t = btk::tuning();
    });
}

// read_namelist.cpp:10-44.  The parse reads key=value lines and stops at the
// first line whose key is not one of the four; the value text of a line
// without '=' is the previous line's (as getline leaves it there).
int btk_read_run_config(const char *path, btk_run_config *cfg) {
    return guarded([&] {
        BTK_REQUIRE(path && cfg, "null argument");
        std::ifstream f(path);
        if (f.fail()) throw btk::Error(std::string("namelist not found: ") + path);
        std::string line, value;
        for (;;) {
            if (!std::getline(f, line)) line.clear();
            std::string key = line;
            const size_t eq = line.find('=');
            if (eq != std::string::npos) {
                key = line.substr(0, eq);
                value = line.substr(eq + 1);
            }
            if (key == "interp_degree") {
                int v = parse_int(value);
                if (v > BTK_MAX_DEGREE) v = BTK_MAX_DEGREE;
                cfg->interp_degree = v;
                cfg->interp_point_count = btk::ipow(v + 1, cfg->dim);
            } else if (key == "cluster_size") {
                cfg->fmm_cluster_thresh = parse_int(value);
                if (cfg->fmm_cluster_thresh <= 0)
                    cfg->fmm_cluster_thresh = btk::ipow(2, cfg->dim) * cfg->interp_point_count;
            } else if (key == "theta") {
                cfg->fmm_theta = parse_double(value);
            } else if (key == "kernel_eps") {
                cfg->ker_eps = parse_double(value);
            } else {
                break;
            }
        }
    });
}

int btk_tree_construction_2d(const btk_run_config *cfg, const double *d_x, const double *d_y, int n, void *stream,
                             btk_tree **tree) {
    return guarded([&] {
        check_cfg(cfg, 2);
        BTK_REQUIRE(tree, "null tree out-pointer");
        auto h = std::make_unique<btk_tree>();
        h->t = btk::build_tree(2, cfg->fmm_cluster_thresh, d_x, d_y, nullptr, n, as_stream(stream));
        *tree = h.release();
    });
}

int btk_tree_construction_3d(const btk_run_config *cfg, const double *d_x, const double *d_y, const double *d_z,
                             int n, void *stream, btk_tree **tree) {
    return guarded([&] {
        check_cfg(cfg, 3);
        BTK_REQUIRE(tree, "null tree out-pointer");
        auto h = std::make_unique<btk_tree>();
        h->t = btk::build_tree(3, cfg->fmm_cluster_thresh, d_x, d_y, d_z, n, as_stream(stream));
        *tree = h.release();
    });
}

int btk_tree_get_info(const btk_tree *tree, btk_tree_info *info) {
    return guarded([&] {
        BTK_REQUIRE(tree && info, "null argument");
        info->panel_count = tree->t->pc;
        info->levels = tree->t->levels;
        info->level_start = tree->t->level_start.data();
    });
}

int btk_tree_copy_panels_2d(const btk_tree *tree, btk_panel_2d *panels) {
    return guarded([&] {
        BTK_REQUIRE(tree && panels, "null argument");
        const btk::Tree &t = *tree->t;
        BTK_REQUIRE(t.dim == 2, "tree is not 2D");
        BTK_HIP(hipDeviceSynchronize());
        std::vector<double> lo[3], hi[3], rad;
        std::vector<int> cnt, par, c0, lvl;
        copy_panels_host(t, lo, hi, rad, cnt, par, c0, lvl);
        for (int p = 0; p < t.pc; p++) {
            btk_panel_2d q;
            std::memset(&q, 0, sizeof q);
            q.id = p;
            q.level = lvl[p];
            q.parent_id = par[p];
            const int b = c0[p];
            q.child1 = b < 0 ? -1 : b;
            q.child2 = b < 0 ? -1 : b + 1;
            q.child3 = b < 0 ? -1 : b + 2;
            q.child4 = b < 0 ? -1 : b + 3;
            q.is_leaf = b < 0;
            q.max_x = hi[0][p];
            q.min_x = lo[0][p];
            q.max_y = hi[1][p];
            q.min_y = lo[1][p];
            q.radius = rad[p];
            q.point_count = cnt[p];
            panels[p] = q;
        }
    });
}

int btk_tree_copy_panels_3d(const btk_tree *tree, btk_panel_3d *panels) {
    return guarded([&] {
        BTK_REQUIRE(tree && panels, "null argument");
        const btk::Tree &t = *tree->t;
        BTK_REQUIRE(t.dim == 3, "tree is not 3D");
        BTK_HIP(hipDeviceSynchronize());
        std::vector<double> lo[3], hi[3], rad;
        std::vector<int> cnt, par, c0, lvl;
        copy_panels_host(t, lo, hi, rad, cnt, par, c0, lvl);
        for (int p = 0; p < t.pc; p++) {
            btk_panel_3d q;
            std::memset(&q, 0, sizeof q);
            q.id = p;
            q.level = lvl[p];
            q.parent_id = par[p];
            const int b = c0[p];
            int *ch[8] = {&q.child1, &q.child2, &q.child3, &q.child4, &q.child5, &q.child6, &q.child7, &q.child8};
            for (int c = 0; c < 8; c++) *ch[c] = b < 0 ? -1 : b + c;
            q.is_leaf = b < 0;
            q.max_x = hi[0][p];
            q.min_x = lo[0][p];
            q.max_y = hi[1][p];
            q.min_y = lo[1][p];
            q.min_z = lo[2][p];
            q.max_z = hi[2][p];
            q.radius = rad[p];
            q.point_count = cnt[p];
            panels[p] = q;
        }
    });
}

int btk_tree_copy_point_leaf_panel(const btk_tree *tree, int *point_leaf_panel) {
    return guarded([&] {
        BTK_REQUIRE(tree && point_leaf_panel, "null argument");
        const btk::Tree &t = *tree->t;
        BTK_HIP(hipDeviceSynchronize());
        if (t.n) BTK_HIP(hipMemcpy(point_leaf_panel, t.point_leaf.p, sizeof(int) * t.n, hipMemcpyDeviceToHost));
    });
}

int btk_tree_copy_permutation(const btk_tree *tree, int *perm, int *start) {
    return guarded([&] {
        BTK_REQUIRE(tree && perm && start, "null argument");
        const btk::Tree &t = *tree->t;
        BTK_HIP(hipDeviceSynchronize());
        if (t.n) BTK_HIP(hipMemcpy(perm, t.perm.p, sizeof(int) * t.n, hipMemcpyDeviceToHost));
        BTK_HIP(hipMemcpy(start, t.start.p, sizeof(int) * t.pc, hipMemcpyDeviceToHost));
        start[t.pc] = t.n;
    });
}

int btk_tree_copy_panel_points_inside(const btk_tree *tree, int width, int *out) {
    return guarded([&] {
        BTK_REQUIRE(tree && out, "null argument");
        BTK_REQUIRE(width > 0, "width must be positive");
        const btk::Tree &t = *tree->t;
        BTK_HIP(hipDeviceSynchronize());
        std::vector<int> perm(t.n), st(t.pc), cnt(t.pc), c0(t.pc);
        if (t.n) BTK_HIP(hipMemcpy(perm.data(), t.perm.p, sizeof(int) * t.n, hipMemcpyDeviceToHost));
        BTK_HIP(hipMemcpy(st.data(), t.start.p, sizeof(int) * t.pc, hipMemcpyDeviceToHost));
        BTK_HIP(hipMemcpy(cnt.data(), t.count.p, sizeof(int) * t.pc, hipMemcpyDeviceToHost));
        BTK_HIP(hipMemcpy(c0.data(), t.child0.p, sizeof(int) * t.pc, hipMemcpyDeviceToHost));
        std::memset(out, 0, sizeof(int) * (size_t)t.pc * width);
        for (int p = 0; p < t.pc; p++) {
            if (c0[p] >= 0) continue;
            BTK_REQUIRE(cnt[p] <= width, "a leaf holds more points than the row width");
            for (int j = 0; j < cnt[p]; j++) out[(size_t)p * width + j] = perm[st[p] + j];
        }
    });
}

void btk_tree_destroy(btk_tree *tree) {
    delete tree;
    btk::end_call();  // no work of the tree's is queued: every call synchronized
}

static int traversal(btk_run_config *cfg, int dim, const btk_tree *target, const btk_tree *source, void *stream,
                     btk_list **list, bool learned_mac = false) {
    return guarded([&] {
        check_cfg(cfg, dim);
        BTK_REQUIRE(target && source && list, "null argument");
        BTK_REQUIRE(target->t->dim == dim && source->t->dim == dim, "tree dimension mismatch");
        auto h = std::make_unique<btk_list>();
        h->l = btk::build_list(cfg, *target->t, *source->t, as_stream(stream), learned_mac);
        *list = h.release();
    });
}

int btk_dual_tree_traversal_2d(btk_run_config *cfg, const btk_tree *target, const btk_tree *source, void *stream,
                               btk_list **list) {
    return traversal(cfg, 2, target, source, stream, list);
}

int btk_dual_tree_traversal_3d(btk_run_config *cfg, const btk_tree *target, const btk_tree *source, void *stream,
                               btk_list **list) {
    return traversal(cfg, 3, target, source, stream, list);
}

int btk_dual_tree_traversal_learned_3d(btk_run_config *cfg, const btk_tree *target, const btk_tree *source,
                                       void *stream, btk_list **list) {
    return traversal(cfg, 3, target, source, stream, list, true);
}

int btk_list_count(const btk_list *list) { return list ? list->l->count : -1; }

int btk_list_copy(const btk_list *list, btk_interact_pair *pairs) {
    return guarded([&] {
        BTK_REQUIRE(list && pairs, "null argument");
        BTK_HIP(hipDeviceSynchronize());
        if (list->l->count)
            BTK_HIP(hipMemcpy(pairs, list->l->pairs.p, sizeof(btk_interact_pair) * list->l->count,
                              hipMemcpyDeviceToHost));
    });
}

// interaction_list.cpp:252-278
int btk_split_interactions(const btk_list *list, btk_interact_pair *pp, btk_interact_pair *pc, btk_interact_pair *cp,
                           btk_interact_pair *cc) {
    return guarded([&] {
        BTK_REQUIRE(list, "null list");
        const btk::List &l = *list->l;
        std::vector<btk_interact_pair> all(l.count);
        BTK_HIP(hipDeviceSynchronize());
        if (l.count)
            BTK_HIP(hipMemcpy(all.data(), l.pairs.p, sizeof(btk_interact_pair) * l.count, hipMemcpyDeviceToHost));
        btk_interact_pair *dst[4] = {pp, pc, cp, cc};
        long long at[4] = {0, 0, 0, 0};
        for (const auto &p : all) {
            BTK_REQUIRE(p.interact_type >= 0 && p.interact_type < 16, "bad interaction type");
            const int ty = p.interact_type & 3;  // bits 2-3: the degree index of a learned-MAC list
            if (dst[ty]) dst[ty][at[ty]] = p;
            at[ty]++;
        }
    });
}

void btk_list_destroy(btk_list *list) {
    delete list;
    btk::end_call();
}

static int upward(const btk_run_config *cfg, int dim, const btk_tree *source, const double *d_charges, double *d_W,
                  void *stream) {
    return guarded([&] {
        check_cfg(cfg, dim);
        BTK_REQUIRE(source && source->t->dim == dim, "bad source tree");
        const btk::Tree &t = *source->t;
        hipStream_t s = as_stream(stream);
        btk::DBuf<double> q(t.n ? t.n : 1);
        btk::gather_sorted(t, d_charges, q.p, s);
        btk::upward_pass(*cfg, t, q.p, d_W, s);
        BTK_HIP(hipStreamSynchronize(s));  // q is released on return
    });
}

// Scratch of btk_upward_pass_3d_noalloc: the source charges in the tree's
// sorted order.  Host code never uses the array's name as an address: each
// device's copy is found once with hipGetSymbolAddress (with that device
// current) and its pointer kept per device.
static __device__ __align__(256) double scratch[BTK_NOALLOC_MAX_POINTS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

double *scratch_on_current_device() {
    int dev = 0;
    BTK_HIP(hipGetDevice(&dev));
    BTK_REQUIRE(dev >= 0 && dev < kMaxDevices, "device id out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        BTK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        BTK_REQUIRE(p != nullptr, "no device address for the scratch array");
        g_scratch[dev] = static_cast<double *>(p);
    }
    return g_scratch[dev];
}
}  // namespace

int btk_upward_pass_3d_noalloc(const btk_run_config *cfg, const btk_tree *source, const double *d_charges,
                               double *d_W, void *stream) {
    return guarded([&] {
        check_cfg(cfg, 3);
        BTK_REQUIRE(source && source->t->dim == 3, "bad source tree");
        const btk::Tree &t = *source->t;
        BTK_REQUIRE(t.n <= BTK_NOALLOC_MAX_POINTS,
                    "btk_upward_pass_3d_noalloc: the source tree has " + std::to_string(t.n) +
                        " points, more than its scratch holds (BTK_NOALLOC_MAX_POINTS = " +
                        std::to_string(BTK_NOALLOC_MAX_POINTS) + ")");
        double *q = scratch_on_current_device();
        hipStream_t s = as_stream(stream);
        btk::gather_sorted(t, d_charges, q, s);
        btk::upward_pass(*cfg, t, q, d_W, s);
    });
}

int btk_upward_pass_2d(const btk_run_config *cfg, const btk_tree *source, const double *d_charges, double *d_W,
                       void *stream) {
    return upward(cfg, 2, source, d_charges, d_W, stream);
}

int btk_upward_pass_3d(const btk_run_config *cfg, const btk_tree *source, const double *d_charges, double *d_W,
                       void *stream) {
    return upward(cfg, 3, source, d_charges, d_W, stream);
}

static int interactions(const btk_run_config *cfg, int dim, const btk_options *opts, const btk_tree *target,
                        const btk_tree *source, const btk_list *list, const double *d_charges, const double *d_W,
                        double *d_soln, double *d_Psi, void *stream) {
    return guarded([&] {
        check_cfg(cfg, dim);
        BTK_REQUIRE(target && source && list, "null argument");
        BTK_REQUIRE(target->t->dim == dim && source->t->dim == dim, "tree dimension mismatch");
        const btk_options o = opts_or_default(opts);
        hipStream_t s = as_stream(stream);
        BTK_REQUIRE(list->l->ndeg == 1,
                    "a learned-MAC list mixes interpolation degrees; its interactions run inside btk_solve");
        const btk::Tree &ts = *source->t;
        btk::DBuf<double> q(ts.n ? ts.n : 1);
        btk::gather_sorted(ts, d_charges, q.p, s);
        btk::fmm_interactions(*cfg, o, *target->t, ts, *list->l, q.p, d_W, d_soln, d_Psi, s);
        BTK_HIP(hipStreamSynchronize(s));  // poisson_fmm.cpp:25 fences
    });
}

int btk_fmm_interactions_2d(const btk_run_config *cfg, const btk_options *opts, const btk_tree *target,
                            const btk_tree *source, const btk_list *list, const double *d_charges, const double *d_W,
                            double *d_soln, double *d_Psi, void *stream) {
    return interactions(cfg, 2, opts, target, source, list, d_charges, d_W, d_soln, d_Psi, stream);
}

int btk_fmm_interactions_3d(const btk_run_config *cfg, const btk_options *opts, const btk_tree *target,
                            const btk_tree *source, const btk_list *list, const double *d_charges, const double *d_W,
                            double *d_soln, double *d_Psi, void *stream) {
    return interactions(cfg, 3, opts, target, source, list, d_charges, d_W, d_soln, d_Psi, stream);
}

static int downward(const btk_run_config *cfg, int dim, const btk_options *opts, const btk_tree *target,
                    double *d_Psi, double *d_soln, void *stream) {
    return guarded([&] {
        check_cfg(cfg, dim);
        BTK_REQUIRE(target && target->t->dim == dim, "bad target tree");
        const btk_options o = opts_or_default(opts);
        btk::downward_pass(*cfg, o, *target->t, d_Psi, d_soln, as_stream(stream));
    });
}

int btk_downward_pass_2d(const btk_run_config *cfg, const btk_options *opts, const btk_tree *target, double *d_Psi,
                         double *d_soln, void *stream) {
    return downward(cfg, 2, opts, target, d_Psi, d_soln, stream);
}

int btk_downward_pass_3d(const btk_run_config *cfg, const btk_options *opts, const btk_tree *target, double *d_Psi,
                         double *d_soln, void *stream) {
    return downward(cfg, 3, opts, target, d_Psi, d_soln, stream);
}

int btk_direct_sum_2d(const btk_options *opts, double eps, int m, const int *d_idx, const double *d_xt,
                      const double *d_yt, int n, const double *d_xs, const double *d_ys, const double *d_q,
                      double *d_out, void *stream) {
    return guarded([&] {
        const btk_options o = opts_or_default(opts);
        btk::direct_sum(2, o, eps, m, d_idx, d_xt, d_yt, nullptr, n, d_xs, d_ys, nullptr, d_q, d_out,
                        as_stream(stream));
    });
}

int btk_direct_sum_3d(const btk_options *opts, double eps, int m, const int *d_idx, const double *d_xt,
                      const double *d_yt, const double *d_zt, int n, const double *d_xs, const double *d_ys,
                      const double *d_zs, const double *d_q, double *d_out, void *stream) {
    return guarded([&] {
        const btk_options o = opts_or_default(opts);
        btk::direct_sum(3, o, eps, m, d_idx, d_xt, d_yt, d_zt, n, d_xs, d_ys, d_zs, d_q, d_out, as_stream(stream));
    });
}

// example_poisson_3d.cpp:70-160 as one call; phase times by host clock with
// the stream drained at every phase boundary.
int btk_solve(const btk_run_config *cfg_in, const btk_options *opts, int m, const double *d_xt, const double *d_yt,
              const double *d_zt, int n, const double *d_xs, const double *d_ys, const double *d_zs,
              const double *d_q, int coincident, double *d_out, btk_solve_stats *stats, void *stream) {
    return guarded([&] {
        BTK_REQUIRE(cfg_in, "null run config");
        btk_run_config cfg = *cfg_in;
        const int dim = cfg.dim;
        BTK_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        const btk_options o = opts_or_default(opts);
        hipStream_t s = as_stream(stream);
        using clk = std::chrono::steady_clock;
        auto sec = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
        const auto t0 = clk::now();
        auto ts = btk::build_tree(dim, cfg.fmm_cluster_thresh, d_xs, d_ys, d_zs, n, s);
        std::unique_ptr<btk::Tree> tt_own;
        if (!coincident) tt_own = btk::build_tree(dim, cfg.fmm_cluster_thresh, d_xt, d_yt, d_zt, m, s);
        const btk::Tree &tt = coincident ? *ts : *tt_own;
        BTK_REQUIRE(!coincident || m == n, "coincident targets and sources must have the same count");
        BTK_HIP(hipStreamSynchronize(s));
        const auto t1 = clk::now();
        auto l = btk::build_list(&cfg, tt, *ts, s, o.mac != 0);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t2 = clk::now();
        const int C = btk::components(dim, o.kernel);
        // the near field and leaf-to-point accumulate in the target tree's
        // sorted order (coalesced adds), moved to the caller's order once
        btk::DBuf<double> q(n ? n : 1), W(btk::proxy_offset(*l, ts->pc, 1, l->ndeg)),
            Psi(btk::proxy_offset(*l, tt.pc, C, l->ndeg)), outs((size_t)C * (m ? m : 1));
        btk::gather_sorted(*ts, d_q, q.p, s);
        btk::upward_all(cfg, *l, *ts, q.p, W.p, s);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t3 = clk::now();
        BTK_HIP(hipMemsetAsync(Psi.p, 0, sizeof(double) * Psi.n, s));
        BTK_HIP(hipMemsetAsync(outs.p, 0, sizeof(double) * outs.n, s));
        btk::KernelEvents kev;
        for (auto &e : kev.ev) BTK_HIP(hipEventCreate(&e));
        btk::fmm_interactions(cfg, o, tt, *ts, *l, q.p, W.p, outs.p, Psi.p, s, &kev, true);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t4 = clk::now();
        float ms_near = 0.f, ms_far = 0.f;
        if (kev.near_done) BTK_HIP(hipEventElapsedTime(&ms_near, kev.ev[0], kev.ev[1]));
        if (kev.far_done) BTK_HIP(hipEventElapsedTime(&ms_far, kev.ev[2], kev.ev[3]));
        for (auto &e : kev.ev) (void)hipEventDestroy(e);
        btk::downward_all(cfg, o, *l, tt, Psi.p, outs.p, s, true);
        btk::add_unsorted(tt, C, outs.p, d_out, s);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t5 = clk::now();
        if (stats) {
            stats->t_tree = sec(t0, t1);
            stats->t_traversal = sec(t1, t2);
            stats->t_upward = sec(t2, t3);
            stats->t_interact = sec(t3, t4);
            stats->t_downward = sec(t4, t5);
            stats->t_total = sec(t0, t5);
            for (int k = 0; k < 4; k++) {
                stats->count[k] = l->counts[k];
                stats->evals[k] = l->evals[k];
            }
            stats->t_kernel_near = 1e-3 * ms_near;
            stats->t_kernel_far = 1e-3 * ms_far;
            stats->cc_shared = kev.cc_shared;
            stats->panel_count_target = tt.pc;
            stats->panel_count_source = ts->pc;
            stats->levels_target = tt.levels;
            stats->levels_source = ts->levels;
        }
    });
}

int btk_autotune(const btk_run_config *cfg_in, const btk_options *opts, int m, const double *d_xt,
                 const double *d_yt, const double *d_zt, int n, const double *d_xs, const double *d_ys,
                 const double *d_zs, const double *d_q, int coincident, btk_tuning *chosen, void *stream) {
    return guarded([&] {
        BTK_REQUIRE(cfg_in, "null run config");
        btk_run_config cfg = *cfg_in;
        const int dim = cfg.dim;
        BTK_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        BTK_REQUIRE(!coincident || m == n, "coincident targets and sources must have the same count");
        const btk_options o = opts_or_default(opts);
        hipStream_t s = as_stream(stream);
        auto ts = btk::build_tree(dim, cfg.fmm_cluster_thresh, d_xs, d_ys, d_zs, n, s);
        std::unique_ptr<btk::Tree> tt_own;
        if (!coincident) tt_own = btk::build_tree(dim, cfg.fmm_cluster_thresh, d_xt, d_yt, d_zt, m, s);
        const btk::Tree &tt = coincident ? *ts : *tt_own;
        auto l = btk::build_list(&cfg, tt, *ts, s, o.mac != 0);
        const int C = btk::components(dim, o.kernel);
        btk::DBuf<double> q(n ? n : 1), W(btk::proxy_offset(*l, ts->pc, 1, l->ndeg)),
            Psi(btk::proxy_offset(*l, tt.pc, C, l->ndeg)), out((size_t)C * (m ? m : 1));
        btk::gather_sorted(*ts, d_q, q.p, s);
        btk::upward_all(cfg, *l, *ts, q.p, W.p, s);
        // Candidate shapes: work-item targets x wavefronts per workgroup x
        // targets per lane x mutual (the last two in mixed precision only).
        // The near and far launches are independent, so call i tries near
        // shape i and far shape i.
        const bool mixed = o.precision == BTK_PREC_MIXED;
        const long long items[] = {8192, 16384, 32768, 65536, 131072};
        const int waves[] = {1, 2, 4};
        struct Shape {
            int items, tpt, mutual, waves;
        };
        std::vector<Shape> nc, fc;
        for (int mu = 0; mu <= (mixed && coincident ? 1 : 0); mu++)
            for (int tpt = 1; tpt <= (mixed ? 2 : 1); tpt++)
                for (int w : waves)
                    for (long long it : items) nc.push_back({(int)it, tpt, mu, w});
        for (int tpt = 1; tpt <= (mixed ? 2 : 1); tpt++)
            for (int w : waves)
                for (long long it : items) fc.push_back({(int)it, tpt, 0, w});
        const int ncall = (int)std::max(nc.size(), fc.size());
        const btk_tuning saved = btk::tuning();
        btk::KernelEvents kev;
        for (auto &e : kev.ev) BTK_HIP(hipEventCreate(&e));
        // each candidate's best time over `passes` passes: 3 when a solve's
        // interactions take under 50 ms (timing noise is then a few percent
        // of a candidate), 2 under 500 ms, else 1; odd passes run the
        // candidates in reverse order, so that a drift of the clocks over a
        // pass favours no candidate
        std::vector<float> tn(nc.size(), 1e30f), tf(fc.size(), 1e30f);
        int passes = 1;
        try {
            for (int pass = 0; pass < passes; pass++) {
                for (int k = pass == 0 ? -1 : 0; k < ncall; k++) {  // k = -1 warms up
                    const int i = (pass & 1) ? ncall - 1 - k : k;
                    const Shape &a = nc[std::min<int>(std::max(i, 0), (int)nc.size() - 1)];
                    const Shape &b = fc[std::min<int>(std::max(i, 0), (int)fc.size() - 1)];
                    btk_tuning t{};
                    t.near_items = a.items;
                    t.near_tpt = a.tpt;
                    t.mutual = a.mutual;
                    t.near_waves = a.waves;
                    t.far_items = b.items;
                    t.far_tpt = b.tpt;
                    t.far_waves = b.waves;
                    t.cc_group = saved.cc_group;
                    t.multi_vectors = saved.multi_vectors;
                    btk::tuning() = t;
                    kev.near_done = kev.far_done = false;
                    const auto c0 = std::chrono::steady_clock::now();
                    btk::fmm_interactions(cfg, o, tt, *ts, *l, q.p, W.p, out.p, Psi.p, s, &kev, true);
                    BTK_HIP(hipStreamSynchronize(s));
                    if (i < 0) {
                        const double sec =
                            std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
                        passes = sec < 0.05 ? 3 : (sec < 0.5 ? 2 : 1);
                        continue;
                    }
                    float ms = 0.f;
                    if (kev.near_done && i < (int)nc.size()) {
                        BTK_HIP(hipEventElapsedTime(&ms, kev.ev[0], kev.ev[1]));
                        tn[i] = std::min(tn[i], ms);
                    }
                    if (kev.far_done && i < (int)fc.size()) {
                        BTK_HIP(hipEventElapsedTime(&ms, kev.ev[2], kev.ev[3]));
                        tf[i] = std::min(tf[i], ms);
                    }
                }
            }
        } catch (...) {
            btk::tuning() = saved;
            for (auto &e : kev.ev) (void)hipEventDestroy(e);
            throw;
        }
        for (auto &e : kev.ev) (void)hipEventDestroy(e);
        btk_tuning best{};
        const size_t in = std::min_element(tn.begin(), tn.end()) - tn.begin();
        const size_t jf = std::min_element(tf.begin(), tf.end()) - tf.begin();
        if (tn[in] < 1e30f) {
            best.near_items = nc[in].items;
            best.near_tpt = nc[in].tpt;
            best.mutual = nc[in].mutual;
            best.near_waves = nc[in].waves;
        }
        if (tf[jf] < 1e30f) {
            best.far_items = fc[jf].items;
            best.far_tpt = fc[jf].tpt;
            best.far_waves = fc[jf].waves;
        }
        best.cc_group = saved.cc_group;
        best.multi_vectors = saved.multi_vectors;
        if (o.precision != BTK_PREC_MIXED) best.near_tpt = best.far_tpt = 0;
        btk::tuning() = best;
        if (chosen) *chosen = best;
    });
}

int btk_autotune_multi(const btk_run_config *cfg_in, const btk_options *opts, int m, const double *d_xt,
                       const double *d_yt, const double *d_zt, int n, const double *d_xs, const double *d_ys,
                       const double *d_zs, const double *d_q, int nrhs, int coincident, btk_tuning *chosen,
                       void *stream) {
    return guarded([&] {
        BTK_REQUIRE(cfg_in, "null run config");
        btk_run_config cfg = *cfg_in;
        const int dim = cfg.dim;
        BTK_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        const btk_options o = opts_or_default(opts);
        BTK_REQUIRE(o.kernel != BTK_KERNEL_REF_VEL3D, "btk_autotune_multi: potential or potential+gradient kernels");
        BTK_REQUIRE(o.precision == BTK_PREC_MIXED && nrhs >= 16,
                    "btk_autotune_multi: the MFMA path needs mixed precision and at least 16 charge vectors");
        BTK_REQUIRE(!coincident || m == n, "coincident targets and sources must have the same count");
        hipStream_t s = as_stream(stream);
        auto ts = btk::build_tree(dim, cfg.fmm_cluster_thresh, d_xs, d_ys, d_zs, n, s);
        std::unique_ptr<btk::Tree> tt_own;
        if (!coincident) tt_own = btk::build_tree(dim, cfg.fmm_cluster_thresh, d_xt, d_yt, d_zt, m, s);
        const btk::Tree &tt = coincident ? *ts : *tt_own;
        auto l = btk::build_list(&cfg, tt, *ts, s, o.mac != 0);
        BTK_REQUIRE(l->ndeg == 1, "btk_autotune_multi: one interpolation degree per list");
        const int C = btk::components(dim, o.kernel);
        const size_t wrows = btk::proxy_offset(*l, ts->pc, 1, l->ndeg), prow = btk::proxy_offset(*l, tt.pc, C, l->ndeg);
        const int kp = (nrhs + 63) / 64 * 64;
        btk::DBuf<float> qs((size_t)(n ? n : 1) * kp), wt(wrows * kp);
        {
            btk::DBuf<double> q(n ? n : 1), W((size_t)nrhs * wrows);
            for (int r = 0; r < nrhs; r++) {
                btk::gather_sorted(*ts, d_q + (size_t)r * n, q.p, s);
                btk::upward_all(cfg, *l, *ts, q.p, W.p + (size_t)r * wrows, s);
            }
            btk::gather_multi_f32(*ts, d_q, nrhs, kp, qs.p, s);
            btk::transpose_weights_f32((long long)wrows, nrhs, kp, W.p, wt.p, s);
            BTK_HIP(hipStreamSynchronize(s));
        }
        btk::DBuf<double> Psi((size_t)nrhs * prow), out((size_t)nrhs * C * (m ? m : 1));
        const int cand[] = {32, 64, 128};
        const int ncand = 3;
        const btk_tuning saved = btk::tuning();
        btk::KernelEvents kev;
        for (auto &e : kev.ev) BTK_HIP(hipEventCreate(&e));
        // best near + far time of each candidate over `passes` passes (3 when
        // the interactions take under 50 ms, 2 under 500 ms, else 1); odd
        // passes run the candidates in reverse order
        float best_ms[ncand];
        for (float &b : best_ms) b = 1e30f;
        int passes = 1;
        try {
            for (int pass = 0; pass < passes; pass++) {
                for (int k = pass == 0 ? -1 : 0; k < ncand; k++) {  // k = -1 warms up
                    const int i = (pass & 1) ? ncand - 1 - k : k;
                    btk_tuning t = saved;
                    t.multi_vectors = cand[std::max(i, 0)];
                    btk::tuning() = t;
                    kev.near_done = kev.far_done = false;
                    const auto c0 = std::chrono::steady_clock::now();
                    btk::fmm_interactions_multi(cfg, o, tt, *ts, *l, qs.p, wt.p, nrhs, kp, out.p, Psi.p, s, &kev);
                    BTK_HIP(hipStreamSynchronize(s));
                    if (i < 0) {
                        const double sec =
                            std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
                        passes = sec < 0.05 ? 3 : (sec < 0.5 ? 2 : 1);
                        continue;
                    }
                    float total = 0.f, ms = 0.f;
                    if (kev.near_done) {
                        BTK_HIP(hipEventElapsedTime(&ms, kev.ev[0], kev.ev[1]));
                        total += ms;
                    }
                    if (kev.far_done) {
                        BTK_HIP(hipEventElapsedTime(&ms, kev.ev[2], kev.ev[3]));
                        total += ms;
                    }
                    best_ms[i] = std::min(best_ms[i], total);
                }
            }
        } catch (...) {
            btk::tuning() = saved;
            for (auto &e : kev.ev) (void)hipEventDestroy(e);
            throw;
        }
        for (auto &e : kev.ev) (void)hipEventDestroy(e);
        btk_tuning best = saved;
        best.multi_vectors = cand[std::min_element(best_ms, best_ms + ncand) - best_ms];
        btk::tuning() = best;
        if (chosen) *chosen = best;
    });
}

int btk_mac_sample(const btk_run_config *cfg, int n, const double *d_x, const double *d_y, const double *d_z,
                   int n_samples, unsigned seed, float *d_X, float *d_Y, int *d_pairs, int *count, void *stream) {
    return guarded([&] {
        BTK_REQUIRE(cfg && count, "null argument");
        *count = btk::mac_sample(*cfg, n, d_x, d_y, d_z, n_samples, seed, d_X, d_Y, d_pairs, as_stream(stream));
    });
}

int btk_mac_train(const float *d_X, const float *d_Y, int n, int epochs, int batch, float lr, unsigned seed,
                  int reinit, float *loss, void *stream) {
    return guarded([&] {
        const float l = btk::mac_train(d_X, d_Y, n, epochs, batch, lr, seed, reinit != 0, as_stream(stream));
        if (loss) *loss = l;
    });
}

int btk_mac_set_weights(const float *h_weights, int count) {
    return guarded([&] {
        BTK_REQUIRE(h_weights && count == BTK_MAC_WEIGHTS, "expected BTK_MAC_WEIGHTS weights");
        btk::mac_set_weights(h_weights, nullptr);
    });
}

int btk_mac_get_weights(float *h_weights, int count) {
    return guarded([&] {
        BTK_REQUIRE(h_weights && count == BTK_MAC_WEIGHTS, "expected BTK_MAC_WEIGHTS weights");
        btk::mac_get_weights(h_weights, nullptr);
    });
}

int btk_mac_set_budget(double rel_err) {
    return guarded([&] {
        BTK_REQUIRE(rel_err > 0 && rel_err < 1, "the budget is a relative error in (0, 1)");
        btk::mac_model().budget = rel_err;
    });
}

int btk_mac_set_costs(const float *costs) {
    return guarded([&] {
        BTK_REQUIRE(costs, "null argument");
        for (int k = 0; k < 4; k++)
            BTK_REQUIRE(costs[k] > 0 && std::isfinite(costs[k]), "costs per evaluation must be positive");
        for (int k = 0; k < 4; k++) btk::mac_model().cost[k] = costs[k];
    });
}

int btk_mac_forward(const float *d_X, int n, float *d_out, void *stream) {
    return guarded([&] {
        btk::mac_forward(d_X, n, d_out, as_stream(stream));
        BTK_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

int btk_mac_gradients(const float *d_X, const float *d_Y, int n, float *d_grad, float *loss, void *stream) {
    return guarded([&] {
        const float l = btk::mac_gradients(d_X, d_Y, n, d_grad, as_stream(stream));
        if (loss) *loss = l;
    });
}

int btk_plan_create(const btk_run_config *cfg_in, const btk_options *opts, int m, const double *d_xt,
                    const double *d_yt, const double *d_zt, int n, const double *d_xs, const double *d_ys,
                    const double *d_zs, int coincident, void *stream, btk_plan **plan) {
    return guarded([&] {
        BTK_REQUIRE(cfg_in && plan, "null argument");
        BTK_REQUIRE(!coincident || m == n, "coincident targets and sources must have the same count");
        auto p = std::make_unique<btk_plan>();
        p->cfg = *cfg_in;
        const int dim = p->cfg.dim;
        BTK_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        p->o = opts_or_default(opts);
        hipStream_t s = as_stream(stream);
        p->ts = btk::build_tree(dim, p->cfg.fmm_cluster_thresh, d_xs, d_ys, d_zs, n, s);
        if (!coincident) p->tt_own = btk::build_tree(dim, p->cfg.fmm_cluster_thresh, d_xt, d_yt, d_zt, m, s);
        p->tt = coincident ? p->ts.get() : p->tt_own.get();
        p->l = btk::build_list(&p->cfg, *p->tt, *p->ts, s, p->o.mac != 0);
        p->ip = btk::plan_interactions(p->cfg, p->o, *p->tt, *p->ts, *p->l, s, true);
        p->m = m;
        p->n = n;
        p->C = btk::components(dim, p->o.kernel);
        p->q.resize(n ? n : 1);
        p->W.resize(btk::proxy_offset(*p->l, p->ts->pc, 1, p->l->ndeg));
        p->Psi.resize(btk::proxy_offset(*p->l, p->tt->pc, p->C, p->l->ndeg));
        p->outs.resize((size_t)p->C * (m ? m : 1));
        BTK_HIP(hipStreamCreateWithFlags(&p->cs, hipStreamNonBlocking));
        BTK_HIP(hipStreamSynchronize(s));
        *plan = p.release();
    });
}

int btk_plan_create_batch(const btk_run_config *cfg_in, const btk_options *opts, int nprob, const int *offsets,
                          const double *d_x, const double *d_y, const double *d_z, void *stream, btk_plan **plan) {
    return guarded([&] {
        BTK_REQUIRE(cfg_in && offsets && plan, "null argument");
        BTK_REQUIRE(nprob >= 1, "nprob must be positive");
        auto p = std::make_unique<btk_plan>();
        p->cfg = *cfg_in;
        const int dim = p->cfg.dim;
        BTK_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        p->o = opts_or_default(opts);
        const std::vector<int> off(offsets, offsets + nprob + 1);
        BTK_REQUIRE(off[0] == 0, "offsets[0] must be 0");
        for (int k = 0; k < nprob; k++) BTK_REQUIRE(off[k + 1] > off[k], "every problem needs at least one point");
        BTK_REQUIRE(!(p->o.deepest_only && nprob > 1), "deepest_only is defined for a single problem only");
        hipStream_t s = as_stream(stream);
        p->ts = btk::build_forest(dim, p->cfg.fmm_cluster_thresh, d_x, d_y, d_z, off, s);
        p->tt = p->ts.get();
        p->l = btk::build_list(&p->cfg, *p->tt, *p->ts, s, p->o.mac != 0);
        p->ip = btk::plan_interactions(p->cfg, p->o, *p->tt, *p->ts, *p->l, s, true);
        p->m = p->n = p->ts->n;
        p->C = btk::components(dim, p->o.kernel);
        p->q.resize(p->n ? p->n : 1);
        p->W.resize(btk::proxy_offset(*p->l, p->ts->pc, 1, p->l->ndeg));
        p->Psi.resize(btk::proxy_offset(*p->l, p->tt->pc, p->C, p->l->ndeg));
        p->outs.resize((size_t)p->C * (p->m ? p->m : 1));
        BTK_HIP(hipStreamCreateWithFlags(&p->cs, hipStreamNonBlocking));
        BTK_HIP(hipStreamSynchronize(s));
        *plan = p.release();
    });
}

int btk_plan_execute(btk_plan *p, const double *d_q, double *d_out, void *stream) {
    return guarded([&] {
        BTK_REQUIRE(p, "null plan");
        hipStream_t s = as_stream(stream);
        if (!p->exec || d_q != p->g_q || d_out != p->g_out) {
            p->drop_graph();
            // launches only: every buffer belongs to the plan or the caller
            BTK_HIP(hipStreamBeginCapture(p->cs, hipStreamCaptureModeThreadLocal));
            try {
                btk::gather_sorted(*p->ts, d_q, p->q.p, p->cs);
                btk::upward_all(p->cfg, *p->l, *p->ts, p->q.p, p->W.p, p->cs);
                btk::zero_fill(p->Psi.p, p->Psi.n, p->cs);
                btk::zero_fill(p->outs.p, p->outs.n, p->cs);
                btk::launch_interactions(*p->ip, p->q.p, p->W.p, p->outs.p, p->Psi.p, p->cs);
                btk::downward_all(p->cfg, p->o, *p->l, *p->tt, p->Psi.p, p->outs.p, p->cs, true);
                btk::add_unsorted(*p->tt, p->C, p->outs.p, d_out, p->cs);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(p->cs, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            BTK_HIP(hipStreamEndCapture(p->cs, &p->graph));
            BTK_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
            p->g_q = d_q;
            p->g_out = d_out;
        }
        BTK_HIP(hipGraphLaunch(p->exec, s));
        BTK_HIP(hipStreamSynchronize(s));
    });
}

void btk_plan_destroy(btk_plan *plan) {
    delete plan;
    btk::end_call();  // btk_plan_execute synchronized, so nothing reads the plan's buffers
}

int btk_solve_batch(const btk_run_config *cfg_in, const btk_options *opts, int nprob, const int *offsets,
                    const double *d_x, const double *d_y, const double *d_z, const double *d_q, double *d_out,
                    btk_solve_stats *stats, void *stream) {
    return guarded([&] {
        BTK_REQUIRE(cfg_in && offsets, "null argument");
        BTK_REQUIRE(nprob >= 1, "nprob must be positive");
        btk_run_config cfg = *cfg_in;
        const int dim = cfg.dim;
        BTK_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        const btk_options o = opts_or_default(opts);
        hipStream_t s = as_stream(stream);
        const std::vector<int> off(offsets, offsets + nprob + 1);
        BTK_REQUIRE(off[0] == 0, "offsets[0] must be 0");
        for (int p = 0; p < nprob; p++) BTK_REQUIRE(off[p + 1] > off[p], "every problem needs at least one point");
        // deepest_only reproduces the reference's level-of-the-last-leaf rule,
        // which is a property of one tree, not of a forest
        BTK_REQUIRE(!(o.deepest_only && nprob > 1), "deepest_only is defined for a single problem only");
        using clk = std::chrono::steady_clock;
        auto sec = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
        const auto t0 = clk::now();
        auto t = btk::build_forest(dim, cfg.fmm_cluster_thresh, d_x, d_y, d_z, off, s);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t1 = clk::now();
        auto l = btk::build_list(&cfg, *t, *t, s, o.mac != 0);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t2 = clk::now();
        const int C = btk::components(dim, o.kernel);
        const int n = t->n;
        btk::DBuf<double> q(n ? n : 1), W(btk::proxy_offset(*l, t->pc, 1, l->ndeg)),
            Psi(btk::proxy_offset(*l, t->pc, C, l->ndeg)), outs((size_t)C * (n ? n : 1));
        btk::gather_sorted(*t, d_q, q.p, s);
        btk::upward_all(cfg, *l, *t, q.p, W.p, s);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t3 = clk::now();
        BTK_HIP(hipMemsetAsync(Psi.p, 0, sizeof(double) * Psi.n, s));
        BTK_HIP(hipMemsetAsync(outs.p, 0, sizeof(double) * outs.n, s));
        btk::KernelEvents kev;
        for (auto &e : kev.ev) BTK_HIP(hipEventCreate(&e));
        btk::fmm_interactions(cfg, o, *t, *t, *l, q.p, W.p, outs.p, Psi.p, s, &kev, true);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t4 = clk::now();
        float ms_near = 0.f, ms_far = 0.f;
        if (kev.near_done) BTK_HIP(hipEventElapsedTime(&ms_near, kev.ev[0], kev.ev[1]));
        if (kev.far_done) BTK_HIP(hipEventElapsedTime(&ms_far, kev.ev[2], kev.ev[3]));
        for (auto &e : kev.ev) (void)hipEventDestroy(e);
        btk::downward_all(cfg, o, *l, *t, Psi.p, outs.p, s, true);
        btk::add_unsorted(*t, C, outs.p, d_out, s);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t5 = clk::now();
        if (stats) {
            stats->t_tree = sec(t0, t1);
            stats->t_traversal = sec(t1, t2);
            stats->t_upward = sec(t2, t3);
            stats->t_interact = sec(t3, t4);
            stats->t_downward = sec(t4, t5);
            stats->t_total = sec(t0, t5);
            for (int k = 0; k < 4; k++) {
                stats->count[k] = l->counts[k];
                stats->evals[k] = l->evals[k];
            }
            stats->t_kernel_near = 1e-3 * ms_near;
            stats->t_kernel_far = 1e-3 * ms_far;
            stats->cc_shared = kev.cc_shared;
            stats->panel_count_target = stats->panel_count_source = t->pc;
            stats->levels_target = stats->levels_source = t->levels;
        }
    });
}

int btk_solve_multi(const btk_run_config *cfg_in, const btk_options *opts, int m, const double *d_xt,
                    const double *d_yt, const double *d_zt, int n, const double *d_xs, const double *d_ys,
                    const double *d_zs, const double *d_q, int nrhs, int coincident, double *d_out,
                    btk_solve_stats *stats, void *stream) {
    return guarded([&] {
        BTK_REQUIRE(cfg_in, "null run config");
        BTK_REQUIRE(nrhs >= 1, "nrhs must be positive");
        btk_run_config cfg = *cfg_in;
        const int dim = cfg.dim;
        BTK_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        const btk_options o = opts_or_default(opts);
        BTK_REQUIRE(o.kernel != BTK_KERNEL_REF_VEL3D, "btk_solve_multi: potential or potential+gradient kernels");
        BTK_REQUIRE(!coincident || m == n, "coincident targets and sources must have the same count");
        hipStream_t s = as_stream(stream);
        using clk = std::chrono::steady_clock;
        auto sec = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
        const auto t0 = clk::now();
        auto ts = btk::build_tree(dim, cfg.fmm_cluster_thresh, d_xs, d_ys, d_zs, n, s);
        std::unique_ptr<btk::Tree> tt_own;
        if (!coincident) tt_own = btk::build_tree(dim, cfg.fmm_cluster_thresh, d_xt, d_yt, d_zt, m, s);
        const btk::Tree &tt = coincident ? *ts : *tt_own;
        BTK_HIP(hipStreamSynchronize(s));
        const auto t1 = clk::now();
        auto l = btk::build_list(&cfg, tt, *ts, s, o.mac != 0);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t2 = clk::now();
        const int C = btk::components(dim, o.kernel);
        const size_t wrows = btk::proxy_offset(*l, ts->pc, 1, l->ndeg), prow = btk::proxy_offset(*l, tt.pc, C, l->ndeg);
        btk::DBuf<double> q((size_t)nrhs * (n ? n : 1)), W((size_t)nrhs * wrows), Psi((size_t)nrhs * prow);
        for (int r = 0; r < nrhs; r++) {
            btk::gather_sorted(*ts, d_q + (size_t)r * n, q.p + (size_t)r * n, s);
            btk::upward_all(cfg, *l, *ts, q.p + (size_t)r * n, W.p + (size_t)r * wrows, s);
        }
        BTK_HIP(hipStreamSynchronize(s));
        const auto t3 = clk::now();
        BTK_HIP(hipMemsetAsync(Psi.p, 0, sizeof(double) * Psi.n, s));
        btk::KernelEvents kev;
        for (auto &e : kev.ev) BTK_HIP(hipEventCreate(&e));
        if (o.precision == BTK_PREC_MIXED && nrhs >= 16 && l->ndeg == 1) {
            const int kp = (nrhs + 63) / 64 * 64;
            btk::DBuf<float> qs((size_t)(n ? n : 1) * kp), wt(wrows * kp);
            btk::gather_multi_f32(*ts, d_q, nrhs, kp, qs.p, s);
            btk::transpose_weights_f32((long long)wrows, nrhs, kp, W.p, wt.p, s);
            btk::fmm_interactions_multi(cfg, o, tt, *ts, *l, qs.p, wt.p, nrhs, kp, d_out, Psi.p, s, &kev);
            BTK_HIP(hipStreamSynchronize(s));
        } else {
            for (int r = 0; r < nrhs; r++)
                btk::fmm_interactions(cfg, o, tt, *ts, *l, q.p + (size_t)r * n, W.p + (size_t)r * wrows,
                                      d_out + (size_t)r * C * m, Psi.p + (size_t)r * prow, s, r == 0 ? &kev : nullptr);
            BTK_HIP(hipStreamSynchronize(s));
        }
        const auto t4 = clk::now();
        float ms_near = 0.f, ms_far = 0.f;
        if (kev.near_done) BTK_HIP(hipEventElapsedTime(&ms_near, kev.ev[0], kev.ev[1]));
        if (kev.far_done) BTK_HIP(hipEventElapsedTime(&ms_far, kev.ev[2], kev.ev[3]));
        for (auto &e : kev.ev) (void)hipEventDestroy(e);
        for (int r = 0; r < nrhs; r++)
            btk::downward_all(cfg, o, *l, tt, Psi.p + (size_t)r * prow, d_out + (size_t)r * C * m, s);
        BTK_HIP(hipStreamSynchronize(s));
        const auto t5 = clk::now();
        if (stats) {
            stats->t_tree = sec(t0, t1);
            stats->t_traversal = sec(t1, t2);
            stats->t_upward = sec(t2, t3);
            stats->t_interact = sec(t3, t4);
            stats->t_downward = sec(t4, t5);
            stats->t_total = sec(t0, t5);
            for (int k = 0; k < 4; k++) {
                stats->count[k] = l->counts[k];
                stats->evals[k] = l->evals[k];
            }
            stats->t_kernel_near = 1e-3 * ms_near;
            stats->t_kernel_far = 1e-3 * ms_far;
            stats->cc_shared = kev.cc_shared;
            stats->panel_count_target = tt.pc;
            stats->panel_count_source = ts->pc;
            stats->levels_target = tt.levels;
            stats->levels_source = ts->levels;
        }
    });
}

}  // extern "C"
