/*
 * btk_oracle.c -- CPU restatement of the BarytreeK treecode (TEST INFRASTRUCTURE).
 *
 * This file is a checker: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  It restates the reference's
 * algorithm loop for loop (citations are to files of the reference
 * repository), in plain C with OpenMP over the same loop extents and
 * atomics that the reference's Kokkos functors use.  Built with
 * -O2 -ffp-contract=off so the geometry (boxes, radii, MAC) rounds exactly as
 * the reference's x86-64 build does.
 *
 * Departures, each deliberate and documented:
 *   - kernels are selected by orc_opts (the reference hard-codes one kernel
 *     per functor); ORC_KERNEL_POT_GRAD has no reference counterpart;
 *   - exclude_self (skip zero-distance pairs) has no reference counterpart;
 *     with it off the reference's self-term (1/eps) is included;
 *   - the downward pass skips leaf panels in parent_to_child: the reference
 *     reads child id -1 there (downward_pass_impl.hpp:20-36), which is
 *     undefined; on trees whose leaves all sit on the deepest level both are
 *     the same computation.
 */
#include "btk_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double ORC_PI = 3.14159265358979323846264338327950288;

void orc_set_threads(int nthreads) {
#ifdef _OPENMP
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
    (void)nthreads;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_components(int dim, int kernel) {
    if (kernel == ORC_KERNEL_POT) return 1;
    if (kernel == ORC_KERNEL_REF_VEL3D) return 3;
    return 1 + dim;
}

static int ipow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; i++) r *= b;
    return r;
}

/* ------------------------------------------------------------------------ */
/* namelist: read_namelist.cpp:10-44                                          */
/* ------------------------------------------------------------------------ */

/* std::stoi semantics: skip leading space, optional sign, >=1 digit. */
static int parse_int(const char *s, int *out) {
    char *end;
    long v = strtol(s, &end, 10);
    if (end == s) return -1;
    *out = (int)v;
    return 0;
}

static int parse_double(const char *s, double *out) {
    char *end;
    double v = strtod(s, &end);
    if (end == s) return -1;
    *out = v;
    return 0;
}

int orc_read_run_config(const char *path, orc_run_config *cfg) {
    FILE *f = fopen(path, "r");
    if (!f) return -1; /* read_namelist.cpp:12-15 throws "namelist not found" */
    char line[4096];
    /* word2 lives across iterations: getline(str1, word2) (:22) leaves it
     * untouched when the line holds no '=' (the stream is already at EOF). */
    char w2[4096] = "";
    for (;;) {
        /* getline(config_file, line) (:19); at EOF the line is taken as empty
         * (the reference would re-read its stale line; here the parse ends) */
        if (!fgets(line, sizeof line, f)) line[0] = '\0';
        size_t len = strlen(line);
        if (len && line[len - 1] == '\n') line[--len] = '\0';
        /* word1 = text up to '=', word2 = rest (:20-22) */
        char *eq = strchr(line, '=');
        if (eq) {
            *eq = '\0';
            snprintf(w2, sizeof w2, "%s", eq + 1);
        }
        const char *w1 = line;
        if (strcmp(w1, "interp_degree") == 0) { /* :24-29 */
            int v;
            if (parse_int(w2, &v)) { fclose(f); return -2; }
            if (v > ORC_MAX_DEGREE) v = ORC_MAX_DEGREE;
            cfg->interp_degree = v;
            cfg->interp_point_count = (int)pow(v + 1, cfg->dim);
        } else if (strcmp(w1, "cluster_size") == 0) { /* :30-35 */
            int v;
            if (parse_int(w2, &v)) { fclose(f); return -2; }
            cfg->fmm_cluster_thresh = v;
            if (v <= 0) cfg->fmm_cluster_thresh = (int)(pow(2, cfg->dim) * cfg->interp_point_count);
        } else if (strcmp(w1, "theta") == 0) { /* :36-37 */
            if (parse_double(w2, &cfg->fmm_theta)) { fclose(f); return -2; }
        } else if (strcmp(w1, "kernel_eps") == 0) { /* :38-39 */
            if (parse_double(w2, &cfg->ker_eps)) { fclose(f); return -2; }
        } else { /* :40-42: first unknown key ends the parse */
            break;
        }
    }
    fclose(f);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* basis: bli_impl.hpp:8-77                                                  */
/* ------------------------------------------------------------------------ */

void orc_interp_vals_bli(double *basis_vals, double x, double min_x, double max_x, int n) {
    double x_range = 0.5 * (max_x - min_x);
    double x_offset = 0.5 * (max_x + min_x);
    double cheb[ORC_MAX_DEGREE + 1], w[ORC_MAX_DEGREE + 1], f[ORC_MAX_DEGREE + 1];
    const double eps = DBL_EPSILON;
    for (int i = 0; i < n + 1; i++) basis_vals[i] = 0;
    /* nodes cos(pi*i/n) (:22-31); degree 0 is not supported by the reference (:20-22) */
    for (int i = 0; i < n + 1; i++) {
        cheb[i] = cos(ORC_PI * i / n) * x_range + x_offset;
        if (i == 0)
            w[i] = 0.5;
        else if (i == n)
            w[i] = 0.5 * pow(-1, i);
        else
            w[i] = pow(-1, i);
    }
    int found = 0;
    for (int i = 0; i < n + 1; i++) { /* coincidence test (:36-45) */
        if (fabs(x - cheb[i]) < eps) {
            found = 1;
            for (int j = 0; j < n + 1; j++) f[j] = 0;
            f[i] = 1;
            break;
        }
    }
    if (!found) { /* second-kind barycentric formula (:46-56) */
        double den = 0;
        for (int i = 0; i < n + 1; i++) {
            double v = w[i] / (x - cheb[i]);
            f[i] = v;
            den += v;
        }
        for (int i = 0; i < n + 1; i++) f[i] /= den;
    }
    for (int i = 0; i < n + 1; i++) basis_vals[i] = f[i];
}

void orc_bli_points_shift(double *pts, double min_x, double max_x, int n) {
    double x_range = 0.5 * (max_x - min_x);
    double x_offset = 0.5 * (max_x + min_x);
    /* note the different node expression cos(pi/n*i) (:74) */
    for (int i = 0; i <= n; i++) pts[i] = cos(ORC_PI / n * i) * x_range + x_offset;
}

/* ------------------------------------------------------------------------ */
/* tree: tree_construction.cpp:5-306                                         */
/* ------------------------------------------------------------------------ */

typedef struct ivec {
    int *a;
    int n, cap;
} ivec;

static void ivec_push(ivec *v, int x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 16;
        v->a = (int *)realloc(v->a, sizeof(int) * (size_t)v->cap);
    }
    v->a[v->n++] = x;
}

int orc_tree_build(int dim, int thresh, int n, const double *x, const double *y, const double *z,
                   orc_tree *tree) {
    const int nc = dim == 3 ? 8 : 4;
    int cap = 64, size = 1;
    orc_panel *P = (orc_panel *)calloc((size_t)cap, sizeof(orc_panel));
    ivec *lists = (ivec *)calloc((size_t)cap, sizeof(ivec));
    int *plp = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));

    /* root (:139-167) */
    P[0].id = 0;
    P[0].level = 0;
    P[0].parent_id = -1;
    for (int c = 0; c < 8; c++) P[0].child[c] = -1;
    P[0].is_leaf = 1;
    P[0].point_count = n;
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double mx[3] = {DBL_MIN, DBL_MIN, DBL_MIN}; /* numeric_limits::min(), F8 */
    const double *co[3] = {x, y, z};
    for (int i = 0; i < n; i++) {
        plp[i] = 0;
        ivec_push(&lists[0], i);
        for (int d = 0; d < dim; d++) {
            double v = co[d][i];
            mn[d] = (v < mn[d]) ? v : mn[d]; /* std::min(mn, v) */
            mx[d] = (mx[d] < v) ? v : mx[d]; /* std::max(mx, v) */
        }
    }
    for (int d = 0; d < 3; d++) {
        P[0].lo[d] = d < dim ? mn[d] : 0.0;
        P[0].hi[d] = d < dim ? mx[d] : 0.0;
    }

    /* child box pattern (:187-194 / :51-54): 1 = upper half [mid,max] */
    static const int hx[8] = {1, 0, 0, 1, 1, 0, 0, 1};
    static const int hy[8] = {1, 1, 0, 0, 1, 1, 0, 0};
    static const int hz[8] = {1, 1, 1, 1, 0, 0, 0, 0};

    for (int i = 0; i < size; i++) { /* (:171) */
        if (P[i].point_count > thresh && P[i].is_leaf) {
            if (P[i].level >= 200) { /* the reference has no depth limit */
                for (int k = 0; k < size; k++) free(lists[k].a);
                free(lists);
                free(P);
                free(plp);
                return -3;
            }
            P[i].is_leaf = 0;
            int base = size, level = P[i].level;
            if (size + nc > cap) {
                int ncap = 2 * cap;
                while (ncap < size + nc) ncap *= 2;
                P = (orc_panel *)realloc(P, sizeof(orc_panel) * (size_t)ncap);
                lists = (ivec *)realloc(lists, sizeof(ivec) * (size_t)ncap);
                memset(lists + cap, 0, sizeof(ivec) * (size_t)(ncap - cap));
                cap = ncap;
            }
            double m1[3], m2[3], m3[3];
            for (int d = 0; d < 3; d++) {
                m1[d] = P[i].lo[d];
                m3[d] = P[i].hi[d];
                m2[d] = 0.5 * (m1[d] + m3[d]);
            }
            for (int c = 0; c < nc; c++) {
                orc_panel *ch = &P[base + c];
                memset(ch, 0, sizeof *ch);
                ch->id = base + c;
                ch->level = level + 1;
                ch->parent_id = i;
                for (int k = 0; k < 8; k++) ch->child[k] = -1;
                ch->is_leaf = 1;
                ch->point_count = 0;
                int h[3] = {hx[c], hy[c], hz[c]};
                for (int d = 0; d < 3; d++) {
                    if (d >= dim) {
                        ch->lo[d] = ch->hi[d] = 0.0;
                    } else if (h[d]) {
                        ch->lo[d] = m2[d];
                        ch->hi[d] = m3[d];
                    } else {
                        ch->lo[d] = m1[d];
                        ch->hi[d] = m2[d];
                    }
                }
                P[i].child[c] = base + c;
            }
            size += nc;
            /* distribute points in list order (:225-274 / :71-98), strict '>' */
            for (int j = 0; j < lists[i].n; j++) {
                int pi = lists[i].a[j];
                double px = x[pi], py = y[pi];
                int which;
                int q2 = (px > m2[0]) ? ((py > m2[1]) ? 0 : 3) : ((py > m2[1]) ? 1 : 2);
                if (dim == 3) {
                    double pz = z[pi];
                    which = (pz > m2[2]) ? q2 : q2 + 4;
                } else {
                    which = q2;
                }
                P[base + which].point_count += 1;
                ivec_push(&lists[base + which], pi);
                plp[pi] = base + which;
            }
        }
    }

    /* finalize (:278-305) */
    int levels = P[size - 1].level + 1;
    int *ls = (int *)malloc(sizeof(int) * (size_t)(levels + 1));
    ls[0] = 0;
    int target_level = 1;
    for (int i = 0; i < size; i++) {
        double ex = P[i].hi[0] - P[i].lo[0];
        double ey = P[i].hi[1] - P[i].lo[1];
        if (dim == 3) {
            double ez = P[i].hi[2] - P[i].lo[2];
            P[i].radius = 0.5 * sqrt(ex * ex + ey * ey + ez * ez);
        } else {
            P[i].radius = 0.5 * sqrt(ex * ex + ey * ey);
        }
        if (P[i].level == target_level) {
            ls[target_level] = i;
            target_level += 1;
        }
    }
    ls[levels] = size;

    int *lstart = (int *)malloc(sizeof(int) * (size_t)(size + 1));
    int *lpts = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int pos = 0;
    for (int i = 0; i < size; i++) {
        lstart[i] = pos;
        if (P[i].is_leaf) {
            for (int j = 0; j < lists[i].n; j++) lpts[pos++] = lists[i].a[j];
        }
        free(lists[i].a);
    }
    lstart[size] = pos;
    free(lists);

    tree->dim = dim;
    tree->panel_count = size;
    tree->levels = levels;
    tree->level_start = ls;
    tree->panels = P;
    tree->point_leaf_panel = plp;
    tree->leaf_start = lstart;
    tree->leaf_points = lpts;
    tree->n_points = n;
    return 0;
}

void orc_tree_free(orc_tree *t) {
    free(t->level_start);
    free(t->panels);
    free(t->point_leaf_panel);
    free(t->leaf_start);
    free(t->leaf_points);
    memset(t, 0, sizeof *t);
}

void orc_pack_panels_2d(const orc_tree *t, orc_panel_2d *out) {
    for (int i = 0; i < t->panel_count; i++) {
        const orc_panel *p = &t->panels[i];
        orc_panel_2d o;
        memset(&o, 0, sizeof o);
        o.id = p->id;
        o.level = p->level;
        o.parent_id = p->parent_id;
        o.child1 = p->child[0];
        o.child2 = p->child[1];
        o.child3 = p->child[2];
        o.child4 = p->child[3];
        o.is_leaf = p->is_leaf != 0;
        o.max_x = p->hi[0];
        o.min_x = p->lo[0];
        o.max_y = p->hi[1];
        o.min_y = p->lo[1];
        o.radius = p->radius;
        o.point_count = p->point_count;
        out[i] = o;
    }
}

void orc_pack_panels_3d(const orc_tree *t, orc_panel_3d *out) {
    for (int i = 0; i < t->panel_count; i++) {
        const orc_panel *p = &t->panels[i];
        orc_panel_3d o;
        memset(&o, 0, sizeof o);
        o.id = p->id;
        o.level = p->level;
        o.parent_id = p->parent_id;
        o.child1 = p->child[0];
        o.child2 = p->child[1];
        o.child3 = p->child[2];
        o.child4 = p->child[3];
        o.child5 = p->child[4];
        o.child6 = p->child[5];
        o.child7 = p->child[6];
        o.child8 = p->child[7];
        o.is_leaf = p->is_leaf != 0;
        o.max_x = p->hi[0];
        o.min_x = p->lo[0];
        o.max_y = p->hi[1];
        o.min_y = p->lo[1];
        o.min_z = p->lo[2];
        o.max_z = p->hi[2];
        o.radius = p->radius;
        o.point_count = p->point_count;
        out[i] = o;
    }
}

void orc_panel_points_inside(const orc_tree *t, int thresh, int *out) {
    memset(out, 0, sizeof(int) * (size_t)t->panel_count * (size_t)thresh);
    for (int i = 0; i < t->panel_count; i++) {
        if (!t->panels[i].is_leaf) continue;
        int s = t->leaf_start[i], e = t->leaf_start[i + 1];
        for (int j = s; j < e && j - s < thresh; j++) out[(size_t)i * thresh + (j - s)] = t->leaf_points[j];
    }
}

/* ------------------------------------------------------------------------ */
/* traversal: interaction_list.cpp:5-278                                      */
/* ------------------------------------------------------------------------ */

/* Children are pushed c1..c8 (:204-238).  ORC_PLANT_PUSH_ORDER is a planted
 * mistake built only by the list tests: the last two children swap places. */
#ifdef ORC_PLANT_PUSH_ORDER
#define ORC_PUSH_CHILD(c, nc) ((c) >= (nc) - 2 ? (2 * (nc) - 3 - (c)) : (c))
#else
#define ORC_PUSH_CHILD(c, nc) (c)
#endif

int orc_dual_tree_traversal(orc_run_config *cfg, const orc_tree *tt, const orc_tree *ts, orc_pair **list_out) {
    const int dim = tt->dim, nc = dim == 3 ? 8 : 4;
    const int thr = cfg->fmm_cluster_thresh;
    ivec qt = {0}, qs = {0};
    int head = 0;
    int lcap = 1024, lcount = 0;
    orc_pair *list = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)lcap);
    int pp = 0, pc = 0, cp = 0, cc = 0;
    ivec_push(&qt, 0);
    ivec_push(&qs, 0);
    while (head < qt.n) { /* (:136) FIFO */
        int it = qt.a[head], is = qs.a[head];
        head++;
        const orc_panel *T = &tt->panels[it], *S = &ts->panels[is];
        double dist;
        {
            double x1 = 0.5 * (T->lo[0] + T->hi[0]);
            double y1 = 0.5 * (T->lo[1] + T->hi[1]);
            double x2 = 0.5 * (S->lo[0] + S->hi[0]);
            double y2 = 0.5 * (S->lo[1] + S->hi[1]);
            double ex = x2 - x1, ey = y2 - y1;
            if (dim == 3) {
                double z1 = 0.5 * (T->lo[2] + T->hi[2]);
                double z2 = 0.5 * (S->lo[2] + S->hi[2]);
                double ez = z2 - z1;
                dist = sqrt(ex * ex + ey * ey + ez * ez);
            } else {
                dist = sqrt(ex * ex + ey * ey);
            }
        }
        double sep = 100;
        if (dist > 0) sep = (T->radius + S->radius) / dist;
        int emit = -1, refine_t = 0, refine_s = 0;
        if (sep < cfg->fmm_theta) { /* (:155-179) */
            int type = 0;
            if (T->point_count > thr) type += 2;
            if (S->point_count > thr) type += 1;
            emit = type;
        } else { /* (:180-239) */
            if (T->is_leaf && S->is_leaf)
                emit = 0;
            else if (T->is_leaf)
                refine_s = 1;
            else if (S->is_leaf)
                refine_t = 1;
#ifdef ORC_PLANT_REFINE_GE /* planted mistake, built only by the list tests: '>=' for '>' */
            else if (T->point_count >= S->point_count)
#else
            else if (T->point_count > S->point_count)
#endif
                refine_t = 1;
            else
                refine_s = 1;
        }
        if (emit >= 0) {
            if (lcount == lcap) {
                lcap *= 2;
                list = (orc_pair *)realloc(list, sizeof(orc_pair) * (size_t)lcap);
            }
            list[lcount].target_panel = it;
            list[lcount].source_panel = is;
            list[lcount].interact_type = emit;
            lcount++;
            if (emit == 0) pp++;
            else if (emit == 1) pc++;
            else if (emit == 2) cp++;
            else cc++;
        } else if (refine_t) {
            for (int c = 0; c < nc; c++) {
                ivec_push(&qs, is);
                ivec_push(&qt, T->child[ORC_PUSH_CHILD(c, nc)]);
            }
        } else if (refine_s) {
            for (int c = 0; c < nc; c++) {
                ivec_push(&qt, it);
                ivec_push(&qs, S->child[ORC_PUSH_CHILD(c, nc)]);
            }
        }
    }
    free(qt.a);
    free(qs.a);
    cfg->fmm_interaction_count = lcount;
    cfg->fmm_pp_count = pp;
    cfg->fmm_pc_count = pc;
    cfg->fmm_cp_count = cp;
    cfg->fmm_cc_count = cc;
    *list_out = list;
    return lcount;
}

void orc_split_interactions(const orc_pair *list, int count, orc_pair *pp, orc_pair *pc, orc_pair *cp,
                            orc_pair *cc) {
    int a = 0, b = 0, c = 0, d = 0; /* (:252-278), stable */
    for (int i = 0; i < count; i++) {
        switch (list[i].interact_type) {
        case 0: pp[a++] = list[i]; break;
        case 1: pc[b++] = list[i]; break;
        case 2: cp[c++] = list[i]; break;
        case 3: cc[d++] = list[i]; break;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* upward pass: upward_pass.cpp:7-29, upward_pass_impl.hpp:9-187              */
/* ------------------------------------------------------------------------ */

#define ATOMIC_ADD(ptr, val)   \
    do {                       \
        double _v = (val);     \
        _Pragma("omp atomic")  \
        *(ptr) += _v;          \
    } while (0)

/* Timing of the last orc_solve's passes (anterpolation, child-to-parent,
 * parent-to-child, leaf-to-point): seconds and the fraction of each pass's
 * units (points, panels, panels, targets) that it processed. */
static double g_pass_seconds[4], g_pass_fraction[4] = {1, 1, 1, 1};
void orc_last_passes(double *seconds, double *fraction) {
    for (int i = 0; i < 4; i++) {
        seconds[i] = g_pass_seconds[i];
        fraction[i] = g_pass_fraction[i];
    }
}
static double now_s(void);

/* Sampling for CPU-baseline timing: unit i is processed when its block of 64
 * consecutive units is every stride-th one, so that concurrent threads still
 * meet on the same outputs as in the full loops (siblings, one target's pairs). */
static inline int sampled(long long i, int stride) { return stride <= 1 || (i / 64) % stride == 0; }

/* Fraction of the panels from lb on (non-leaves only if parents, weighted by
 * child count; else all) that a pass with this stride processes. */
static double sampled_fraction(const orc_tree *t, int lb, int stride, int parents) {
    double all = 0, done = 0;
    const int nc = t->dim == 3 ? 8 : 4;
    for (int i = lb; i < t->level_start[t->levels]; i++) {
        double w = parents ? (t->panels[i].is_leaf ? 0.0 : (double)nc) : 1.0;
        all += w;
        if (sampled(i, stride)) done += w;
    }
    return all > 0 ? done / all : 1.0;
}

/* stride > 1 (CPU-baseline sampling): only every stride-th point and panel */
static void upward_pass(const orc_run_config *cfg, const orc_tree *ts, const double *x, const double *y,
                        const double *z, const double *q, double *W, int stride) {
    const int dim = ts->dim, n = cfg->interp_degree, n1 = n + 1;
    const int P = ipow(n1, dim);
    const int N = ts->n_points;
    double t0 = now_s();
    /* base_pots (upward_pass_impl.hpp:51-71, :133-159) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        if (!sampled(i, stride)) continue;
        int leaf = ts->point_leaf_panel[i];
        const orc_panel *L = &ts->panels[leaf];
        double bx[ORC_MAX_DEGREE + 1], by[ORC_MAX_DEGREE + 1], bz[ORC_MAX_DEGREE + 1];
        orc_interp_vals_bli(bx, x[i], L->lo[0], L->hi[0], n);
        orc_interp_vals_bli(by, y[i], L->lo[1], L->hi[1], n);
        double *w = W + (size_t)leaf * P;
        if (dim == 3) {
            orc_interp_vals_bli(bz, z[i], L->lo[2], L->hi[2], n);
            int idx = 0;
            for (int j = 0; j < n1; j++)
                for (int k = 0; k < n1; k++)
                    for (int l = 0; l < n1; l++) {
                        ATOMIC_ADD(&w[idx], bx[j] * by[k] * bz[l] * q[i]);
                        idx++;
                    }
        } else {
            int idx = 0;
            for (int j = 0; j < n1; j++)
                for (int k = 0; k < n1; k++) {
                    ATOMIC_ADD(&w[idx], bx[j] * by[k] * q[i]);
                    idx++;
                }
        }
    }
    g_pass_seconds[0] = now_s() - t0;
    long long nd = 0;
    for (int i = 0; i < N; i++) nd += sampled(i, stride);
    g_pass_fraction[0] = N > 0 ? (double)nd / N : 1.0;
    t0 = now_s();
    /* child_to_parent per level, deepest to 1 (upward_pass.cpp:23-28) */
    for (int lev = ts->levels - 1; lev > 0; lev--) {
        int lb = ts->level_start[lev], ub = ts->level_start[lev + 1];
#pragma omp parallel for schedule(dynamic)
        for (int i = lb; i < ub; i++) {
            if (!sampled(i, stride)) continue;
            const orc_panel *C = &ts->panels[i];
            int parent = C->parent_id;
            if (parent <= -1) continue;
            const orc_panel *Pp = &ts->panels[parent];
            double cx[ORC_MAX_DEGREE + 1], cy[ORC_MAX_DEGREE + 1], cz[ORC_MAX_DEGREE + 1];
            double bx[ORC_MAX_DEGREE + 1], by[ORC_MAX_DEGREE + 1], bz[ORC_MAX_DEGREE + 1];
            orc_bli_points_shift(cx, C->lo[0], C->hi[0], n);
            orc_bli_points_shift(cy, C->lo[1], C->hi[1], n);
            const double *wc = W + (size_t)i * P;
            double *wp = W + (size_t)parent * P;
            int idx1 = 0;
            if (dim == 3) {
                orc_bli_points_shift(cz, C->lo[2], C->hi[2], n);
                for (int j1 = 0; j1 < n1; j1++) {
                    orc_interp_vals_bli(bx, cx[j1], Pp->lo[0], Pp->hi[0], n);
                    for (int k1 = 0; k1 < n1; k1++) {
                        orc_interp_vals_bli(by, cy[k1], Pp->lo[1], Pp->hi[1], n);
                        for (int l1 = 0; l1 < n1; l1++) {
                            orc_interp_vals_bli(bz, cz[l1], Pp->lo[2], Pp->hi[2], n);
                            int idx2 = 0;
                            for (int j2 = 0; j2 < n1; j2++)
                                for (int k2 = 0; k2 < n1; k2++)
                                    for (int l2 = 0; l2 < n1; l2++) {
                                        ATOMIC_ADD(&wp[idx2], bx[j2] * by[k2] * bz[l2] * wc[idx1]);
                                        idx2++;
                                    }
                            idx1++;
                        }
                    }
                }
            } else {
                for (int j1 = 0; j1 < n1; j1++) {
                    orc_interp_vals_bli(bx, cx[j1], Pp->lo[0], Pp->hi[0], n);
                    for (int k1 = 0; k1 < n1; k1++) {
                        orc_interp_vals_bli(by, cy[k1], Pp->lo[1], Pp->hi[1], n);
                        int idx2 = 0;
                        for (int j2 = 0; j2 < n1; j2++)
                            for (int k2 = 0; k2 < n1; k2++) {
                                ATOMIC_ADD(&wp[idx2], bx[j2] * by[k2] * wc[idx1]);
                                idx2++;
                            }
                        idx1++;
                    }
                }
            }
        }
    }
    g_pass_seconds[1] = now_s() - t0;
    g_pass_fraction[1] = sampled_fraction(ts, ts->levels > 1 ? ts->level_start[1] : 0, stride, 0);
}

void orc_upward_pass(const orc_run_config *cfg, const orc_tree *ts, const double *x, const double *y,
                     const double *z, const double *q, double *W) {
    upward_pass(cfg, ts, x, y, z, q, W, 1);
}

/* ------------------------------------------------------------------------ */
/* kernel evaluation: poisson_fmm_impl.hpp:38-48,241-254;                    */
/* poisson_vel_fmm_impl.hpp:42-56                                            */
/* ------------------------------------------------------------------------ */

typedef struct kparams {
    int dim, kernel, excl;
    double gfc, eps;
} kparams;

static kparams make_kparams(int dim, double eps, const orc_opts *opt) {
    kparams k;
    k.dim = dim;
    k.kernel = opt->kernel;
    k.excl = opt->exclude_self;
    k.eps = eps;
    double g = dim == 2 ? 1.0 / (4.0 * ORC_PI) : -1.0 / (4.0 * ORC_PI);
    k.gfc = opt->convention ? -g : g;
    return k;
}

/* acc[c] += G_c(t, s) * q */
static inline void keval(const kparams *k, double tx, double ty, double tz, double sx, double sy, double sz,
                         double q, double *acc) {
    double dx = tx - sx, dy = ty - sy;
    if (k->dim == 3) {
        double dz = tz - sz;
        if (k->excl && dx == 0 && dy == 0 && dz == 0) return;
        if (k->kernel == ORC_KERNEL_POT) {
            double gfv = k->gfc / (sqrt(dx * dx + dy * dy + dz * dz + k->eps * k->eps));
            acc[0] += gfv * q;
        } else if (k->kernel == ORC_KERNEL_REF_VEL3D) {
            double gfv = q * k->gfc / (dx * dx + dy * dy + dz * dz + k->eps * k->eps);
            acc[0] += dx * gfv;
            acc[1] += dy * gfv;
            acc[2] += dz * gfv;
        } else {
            double R = dx * dx + dy * dy + dz * dz + k->eps * k->eps;
            double inv = 1.0 / sqrt(R);
            double g = k->gfc * inv;
            acc[0] += g * q;
            double h = -g * inv * inv * q; /* d/dt (gfc/|t-s|) = -gfc (t-s)/|t-s|^3 */
            acc[1] += h * dx;
            acc[2] += h * dy;
            acc[3] += h * dz;
        }
    } else {
        if (k->excl && dx == 0 && dy == 0) return;
        double R = dx * dx + dy * dy + k->eps * k->eps;
        double gfv = k->gfc * log(R);
        acc[0] += gfv * q;
        if (k->kernel == ORC_KERNEL_POT_GRAD) {
            double h = 2.0 * k->gfc * q / R; /* d/dt (gfc log R) = 2 gfc (t-s)/R */
            acc[1] += h * dx;
            acc[2] += h * dy;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* interactions: poisson_fmm.cpp:8-26 + poisson_fmm_impl.hpp:10-442          */
/* ------------------------------------------------------------------------ */

static void proxy_grid(const orc_panel *p, int dim, int n, double *cx, double *cy, double *cz) {
    orc_bli_points_shift(cx, p->lo[0], p->hi[0], n);
    orc_bli_points_shift(cy, p->lo[1], p->hi[1], n);
    if (dim == 3) orc_bli_points_shift(cz, p->lo[2], p->hi[2], n);
}

/* coordinates of proxy point idx of a grid */
static inline void proxy_point(int dim, int n1, int idx, const double *cx, const double *cy, const double *cz,
                               double *px, double *py, double *pz) {
    if (dim == 3) {
        int l = idx % n1, k = (idx / n1) % n1, j = idx / (n1 * n1);
        *px = cx[j];
        *py = cy[k];
        *pz = cz[l];
    } else {
        int k = idx % n1, j = idx / n1;
        *px = cx[j];
        *py = cy[k];
        *pz = 0;
    }
}

static double g_near_seconds = 0, g_near_fraction = 1, g_far_seconds = 0, g_far_fraction = 1;
double orc_last_near_seconds(void) { return g_near_seconds; }
double orc_last_near_fraction(void) { return g_near_fraction; }
double orc_last_far_seconds(void) { return g_far_seconds; }
double orc_last_far_fraction(void) { return g_far_fraction; }

void orc_fmm_interactions(const orc_run_config *cfg, const orc_opts *opt, const orc_tree *tt, const orc_tree *ts,
                          const double *xt, const double *yt, const double *zt, const double *xs,
                          const double *ys, const double *zs, const double *q, const double *W,
                          const orc_pair *list, int count, double *out, double *Psi) {
    const int dim = tt->dim, n = cfg->interp_degree, n1 = n + 1;
    const int P = ipow(n1, dim);
    const int C = orc_components(dim, opt->kernel);
    const int M = tt->n_points;
    const size_t pcT = (size_t)tt->panel_count;
    const kparams kp = make_kparams(dim, cfg->ker_eps, opt);
    static const double zero3[1] = {0};
    /* sampling strides of the near (PP, PC) and far (CP, CC) pairs */
    const int stride[2] = {opt->near_stride > 1 ? opt->near_stride : 1, opt->far_stride > 1 ? opt->far_stride : 1};
    double all[2] = {0, 0}, done[2] = {0, 0};
    for (int i = 0; i < count; i++) {
        int ty = list[i].interact_type, far = ty > 1;
        int it = list[i].target_panel, is = list[i].source_panel;
        double nt = (double)(tt->leaf_start[it + 1] - tt->leaf_start[it]);
        double ns = (double)(ts->leaf_start[is + 1] - ts->leaf_start[is]);
        double e = (ty == 0 || ty == 1 ? nt : (double)P) * (ty == 0 || ty == 2 ? ns : (double)P);
        all[far] += e;
        if (sampled(i, stride[far])) done[far] += e;
    }
    g_near_fraction = all[0] > 0 ? done[0] / all[0] : 1.0;
    g_far_fraction = all[1] > 0 ? done[1] / all[1] : 1.0;
    double tn0 = 0;

    /* one launch per type, PP, PC, CP, CC (poisson_fmm.cpp:21-24) */
    for (int type = 0; type < 4; type++) {
        if (type == 0) tn0 = now_s();
        if (type == 2) {
            g_near_seconds = now_s() - tn0;
            tn0 = now_s();
        }
        const int st = stride[type > 1];
#pragma omp parallel for schedule(dynamic)
        for (int i = 0; i < count; i++) {
            if (list[i].interact_type != type) continue;
            if (!sampled(i, st)) continue;
            const int it = list[i].target_panel, is = list[i].source_panel;
            const orc_panel *T = &tt->panels[it], *S = &ts->panels[is];
            double tcx[ORC_MAX_DEGREE + 1], tcy[ORC_MAX_DEGREE + 1], tcz[ORC_MAX_DEGREE + 1];
            double scx[ORC_MAX_DEGREE + 1], scy[ORC_MAX_DEGREE + 1], scz[ORC_MAX_DEGREE + 1];
            if (type == 0) { /* PP (:235-256): targets outer, sources inner */
                for (int a = tt->leaf_start[it]; a < tt->leaf_start[it + 1]; a++) {
                    int i_t = tt->leaf_points[a];
                    double tz = dim == 3 ? zt[i_t] : 0;
                    for (int b = ts->leaf_start[is]; b < ts->leaf_start[is + 1]; b++) {
                        int i_s = ts->leaf_points[b];
                        double acc[4] = {0, 0, 0, 0};
                        keval(&kp, xt[i_t], yt[i_t], tz, xs[i_s], ys[i_s], dim == 3 ? zs[i_s] : zero3[0], q[i_s], acc);
                        for (int c = 0; c < C; c++) ATOMIC_ADD(&out[(size_t)c * M + i_t], acc[c]);
                    }
                }
            } else if (type == 1) { /* PC (:280-315) */
                proxy_grid(S, dim, n, scx, scy, scz);
                const double *w = W + (size_t)is * P;
                for (int a = tt->leaf_start[it]; a < tt->leaf_start[it + 1]; a++) {
                    int i_t = tt->leaf_points[a];
                    double tz = dim == 3 ? zt[i_t] : 0;
                    for (int idx = 0; idx < P; idx++) {
                        double sx, sy, sz;
                        proxy_point(dim, n1, idx, scx, scy, scz, &sx, &sy, &sz);
                        double acc[4] = {0, 0, 0, 0};
                        keval(&kp, xt[i_t], yt[i_t], tz, sx, sy, sz, w[idx], acc);
                        for (int c = 0; c < C; c++) ATOMIC_ADD(&out[(size_t)c * M + i_t], acc[c]);
                    }
                }
            } else if (type == 2) { /* CP (:339-373): target proxies outer, sources inner */
                proxy_grid(T, dim, n, tcx, tcy, tcz);
                for (int idx = 0; idx < P; idx++) {
                    double tx, ty, tz;
                    proxy_point(dim, n1, idx, tcx, tcy, tcz, &tx, &ty, &tz);
                    for (int b = ts->leaf_start[is]; b < ts->leaf_start[is + 1]; b++) {
                        int i_s = ts->leaf_points[b];
                        double acc[4] = {0, 0, 0, 0};
                        keval(&kp, tx, ty, tz, xs[i_s], ys[i_s], dim == 3 ? zs[i_s] : 0, q[i_s], acc);
                        for (int c = 0; c < C; c++) ATOMIC_ADD(&Psi[((size_t)c * pcT + it) * P + idx], acc[c]);
                    }
                }
            } else { /* CC (:392-441) */
                proxy_grid(T, dim, n, tcx, tcy, tcz);
                proxy_grid(S, dim, n, scx, scy, scz);
                const double *w = W + (size_t)is * P;
                for (int idx = 0; idx < P; idx++) {
                    double tx, ty, tz;
                    proxy_point(dim, n1, idx, tcx, tcy, tcz, &tx, &ty, &tz);
                    for (int sidx = 0; sidx < P; sidx++) {
                        double sx, sy, sz;
                        proxy_point(dim, n1, sidx, scx, scy, scz, &sx, &sy, &sz);
                        double acc[4] = {0, 0, 0, 0};
                        keval(&kp, tx, ty, tz, sx, sy, sz, w[sidx], acc);
                        for (int c = 0; c < C; c++) ATOMIC_ADD(&Psi[((size_t)c * pcT + it) * P + idx], acc[c]);
                    }
                }
            }
        }
    }
    g_far_seconds = now_s() - tn0;
}

/* ------------------------------------------------------------------------ */
/* downward pass: downward_pass.cpp:7-41, downward_pass_impl.hpp:9-379        */
/* ------------------------------------------------------------------------ */

void orc_downward_pass(const orc_run_config *cfg, const orc_opts *opt, const orc_tree *tt, const double *x,
                       const double *y, const double *z, double *Psi, double *out) {
    const int dim = tt->dim, n = cfg->interp_degree, n1 = n + 1;
    const int nc = dim == 3 ? 8 : 4;
    const int P = ipow(n1, dim);
    const int C = orc_components(dim, opt->kernel);
    const int M = tt->n_points;
    const size_t pc = (size_t)tt->panel_count;
    /* >1 (CPU-baseline sampling): only every stride-th panel and leaf */
    const int stride = opt->pass_stride > 1 ? opt->pass_stride : 1;
    double t0 = now_s();
    /* parent_to_child, levels 0..L-2 (downward_pass.cpp:21-25) */
    for (int lev = 0; lev < tt->levels - 1; lev++) {
        int lb = tt->level_start[lev], ub = tt->level_start[lev + 1];
#pragma omp parallel for collapse(2) schedule(dynamic)
        for (int i = lb; i < ub; i++) {
            for (int cj = 0; cj < nc; cj++) {
                const orc_panel *Pp = &tt->panels[i];
                if (Pp->is_leaf) continue; /* reference reads child -1 here (undefined) */
                if (!sampled(i, stride)) continue;
                int child = Pp->child[cj];
                const orc_panel *Cc = &tt->panels[child];
                double cx[ORC_MAX_DEGREE + 1], cy[ORC_MAX_DEGREE + 1], cz[ORC_MAX_DEGREE + 1];
                double bx[ORC_MAX_DEGREE + 1], by[ORC_MAX_DEGREE + 1], bz[ORC_MAX_DEGREE + 1];
                orc_bli_points_shift(cx, Cc->lo[0], Cc->hi[0], n);
                orc_bli_points_shift(cy, Cc->lo[1], Cc->hi[1], n);
                if (dim == 3) orc_bli_points_shift(cz, Cc->lo[2], Cc->hi[2], n);
                for (int comp = 0; comp < C; comp++) {
                    double *pch = Psi + ((size_t)comp * pc + child) * P;
                    const double *ppa = Psi + ((size_t)comp * pc + i) * P;
                    int idx1 = 0;
                    if (dim == 3) {
                        for (int j = 0; j < n1; j++) {
                            orc_interp_vals_bli(bx, cx[j], Pp->lo[0], Pp->hi[0], n);
                            for (int k = 0; k < n1; k++) {
                                orc_interp_vals_bli(by, cy[k], Pp->lo[1], Pp->hi[1], n);
                                for (int l = 0; l < n1; l++) {
                                    orc_interp_vals_bli(bz, cz[l], Pp->lo[2], Pp->hi[2], n);
                                    int idx2 = 0;
                                    for (int j2 = 0; j2 < n1; j2++)
                                        for (int k2 = 0; k2 < n1; k2++)
                                            for (int l2 = 0; l2 < n1; l2++) {
                                                ATOMIC_ADD(&pch[idx1], bx[j2] * by[k2] * bz[l2] * ppa[idx2]);
                                                idx2++;
                                            }
                                    idx1++;
                                }
                            }
                        }
                    } else {
                        for (int j = 0; j < n1; j++) {
                            for (int k = 0; k < n1; k++) {
                                orc_interp_vals_bli(bx, cx[j], Pp->lo[0], Pp->hi[0], n);
                                orc_interp_vals_bli(by, cy[k], Pp->lo[1], Pp->hi[1], n);
                                int idx2 = 0;
                                for (int j2 = 0; j2 < n1; j2++)
                                    for (int k2 = 0; k2 < n1; k2++) {
                                        ATOMIC_ADD(&pch[idx1], bx[j2] * by[k2] * ppa[idx2]);
                                        idx2++;
                                    }
                                idx1++;
                            }
                        }
                    }
                }
            }
        }
    }
    /* leaf_to_point (downward_pass.cpp:26-28): deepest level only in the
     * reference; every leaf in the corrected form */
    g_pass_seconds[2] = now_s() - t0;
    g_pass_fraction[2] = sampled_fraction(tt, 0, stride, 1);
    t0 = now_s();
    int lb = opt->deepest_only ? tt->level_start[tt->levels - 1] : 0;
    int ub = tt->level_start[tt->levels];
    double all = 0, done = 0;
    for (int i = lb; i < ub; i++) {
        if (!tt->panels[i].is_leaf) continue;
        double np = (double)(tt->leaf_start[i + 1] - tt->leaf_start[i]);
        all += np;
        if (sampled(i, stride)) done += np;
    }
#pragma omp parallel for schedule(dynamic)
    for (int i = lb; i < ub; i++) {
        const orc_panel *L = &tt->panels[i];
        if (!L->is_leaf) continue;
        if (!sampled(i, stride)) continue;
        for (int a = tt->leaf_start[i]; a < tt->leaf_start[i + 1]; a++) {
            int i_t = tt->leaf_points[a];
            double bx[ORC_MAX_DEGREE + 1], by[ORC_MAX_DEGREE + 1], bz[ORC_MAX_DEGREE + 1];
            orc_interp_vals_bli(bx, x[i_t], L->lo[0], L->hi[0], n);
            orc_interp_vals_bli(by, y[i_t], L->lo[1], L->hi[1], n);
            if (dim == 3) orc_interp_vals_bli(bz, z[i_t], L->lo[2], L->hi[2], n);
            for (int comp = 0; comp < C; comp++) {
                const double *psi = Psi + ((size_t)comp * pc + i) * P;
                double *o = &out[(size_t)comp * M + i_t];
                int idx = 0;
                if (dim == 3) {
                    for (int k = 0; k < n1; k++)
                        for (int l = 0; l < n1; l++)
                            for (int m = 0; m < n1; m++) {
                                ATOMIC_ADD(o, psi[idx] * bx[k] * by[l] * bz[m]);
                                idx++;
                            }
                } else {
                    for (int k = 0; k < n1; k++)
                        for (int l = 0; l < n1; l++) {
                            ATOMIC_ADD(o, psi[idx] * bx[k] * by[l]);
                            idx++;
                        }
                }
            }
        }
    }
    g_pass_seconds[3] = now_s() - t0;
    g_pass_fraction[3] = all > 0 ? done / all : 1.0;
}

/* ------------------------------------------------------------------------ */
/* direct sum: direct_sum_impl.hpp:9-94                                       */
/* ------------------------------------------------------------------------ */

void orc_direct_sum(int dim, double eps, const orc_opts *opt, int m, const int *idx, const double *xt,
                    const double *yt, const double *zt, int n, const double *xs, const double *ys,
                    const double *zs, const double *q, double *out) {
    const kparams kp = make_kparams(dim, eps, opt);
    const int C = orc_components(dim, opt->kernel);
#pragma omp parallel for schedule(dynamic, 16)
    for (int a = 0; a < m; a++) {
        int i = idx ? idx[a] : a;
        double acc[4] = {0, 0, 0, 0};
        double tz = dim == 3 ? zt[i] : 0;
        for (int j = 0; j < n; j++) keval(&kp, xt[i], yt[i], tz, xs[j], ys[j], dim == 3 ? zs[j] : 0, q[j], acc);
        for (int c = 0; c < C; c++) out[(size_t)c * m + a] += acc[c];
    }
}

/* ------------------------------------------------------------------------ */
/* whole pipeline: example_poisson_3d.cpp:70-160                              */
/* ------------------------------------------------------------------------ */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_solve(const orc_run_config *cfg_in, const orc_opts *opt, int m, const double *xt, const double *yt,
              const double *zt, int n, const double *xs, const double *ys, const double *zs, const double *q,
              int coincident, double *out, double *t, int *counts) {
    orc_run_config cfg = *cfg_in;
    const int dim = cfg.dim;
    const int C = orc_components(dim, opt->kernel);
    double t0 = now_s();
    orc_tree ts, tt;
    int rc = orc_tree_build(dim, cfg.fmm_cluster_thresh, n, xs, ys, zs, &ts);
    if (rc) return rc;
    if (!coincident) {
        rc = orc_tree_build(dim, cfg.fmm_cluster_thresh, m, xt, yt, zt, &tt);
        if (rc) { orc_tree_free(&ts); return rc; }
    }
    const orc_tree *T = coincident ? &ts : &tt;
    double t1 = now_s();
    orc_pair *list = NULL;
    int count = orc_dual_tree_traversal(&cfg, T, &ts, &list);
    double t2 = now_s();
    const int P = ipow(cfg.interp_degree + 1, dim);
    double *W = (double *)calloc((size_t)ts.panel_count * P, sizeof(double));
    double *Psi = (double *)calloc((size_t)C * T->panel_count * P, sizeof(double));
    upward_pass(&cfg, &ts, xs, ys, zs, q, W, opt->pass_stride > 1 ? opt->pass_stride : 1);
    double t3 = now_s();
    orc_fmm_interactions(&cfg, opt, T, &ts, xt, yt, zt, xs, ys, zs, q, W, list, count, out, Psi);
    double t4 = now_s();
    orc_downward_pass(&cfg, opt, T, xt, yt, zt, Psi, out);
    double t5 = now_s();
    if (t) {
        t[0] = t1 - t0;
        t[1] = t2 - t1;
        t[2] = t3 - t2;
        t[3] = t4 - t3;
        t[4] = t5 - t4;
    }
    if (counts) {
        counts[0] = cfg.fmm_pp_count;
        counts[1] = cfg.fmm_pc_count;
        counts[2] = cfg.fmm_cp_count;
        counts[3] = cfg.fmm_cc_count;
    }
    free(W);
    free(Psi);
    free(list);
    if (!coincident) orc_tree_free(&tt);
    orc_tree_free(&ts);
    return 0;
}
