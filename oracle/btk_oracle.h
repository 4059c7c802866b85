/*
 * btk_oracle.h -- CPU restatement of the BarytreeK barycentric Lagrange
 * dual-tree treecode, used ONLY as a test checker.
 *
 * TEST INFRASTRUCTURE.  Nothing in the product (barytreek_amd/, include/)
 * links or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, as the thing results are checked
 * against (or, for the CPU baseline, timed beside the GPU path).
 *
 * Every function restates the reference's algorithm in plain C, following
 * the reference file:line given in its comment.  The reference itself cannot
 * be compiled in this image (it needs Kokkos and a CMake-generated config
 * header), so the restatement is pinned by the checks listed in
 * tests/test_oracle_*.py: closed forms of the barycentric basis, tree and
 * list invariants, hand-derived tiny trees under tests/golden/, brute-force
 * coverage of every (target, source) point pair, and direct-sum limits.
 *
 * Conventions (all arrays are plain C arrays, SoA, FP64):
 *   - panels are returned in a generic form (orc_panel) and can be packed
 *     into the reference's AoS byte layout (TreePanel_2d / TreePanel_3d,
 *     structs.hpp:29-67) with orc_pack_panels_{2d,3d};
 *   - proxy arrays are [panel][P] row-major, P = (n+1)^dim, with flat index
 *     idx = (j*(n+1)+k)*(n+1)+l (upward_pass_impl.hpp:150-156);
 *   - multi-component outputs are component-major: out[c*M + i].
 */
#ifndef BTK_ORACLE_H
#define BTK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_DEGREE 10 /* barytreek-config.h.in:20 */

/* Byte-for-byte mirror of RunConfig (structs.hpp:6-21). */
typedef struct orc_run_config {
    int dim;
    int interp_degree;
    int interp_point_count;
    double fmm_theta;
    int fmm_cluster_thresh;
    int fmm_interaction_count;
    int fmm_pp_count;
    int fmm_pc_count;
    int fmm_cp_count;
    int fmm_cc_count;
    double ker_eps;
    int mpi_p;
    int mpi_id;
} orc_run_config;

/* Generic (dimension-agnostic) panel used inside the oracle. */
typedef struct orc_panel {
    int id, level, parent_id;
    int child[8];      /* -1 when absent; 2D uses child[0..3] */
    int is_leaf;
    double lo[3], hi[3];
    double radius;
    int point_count;
} orc_panel;

/* Byte-for-byte mirrors of TreePanel_2d / TreePanel_3d (structs.hpp:29-67). */
typedef struct orc_panel_2d {
    int id, level, parent_id, child1, child2, child3, child4;
    _Bool is_leaf;
    double max_x, min_x, max_y, min_y, radius;
    int point_count;
} orc_panel_2d;

typedef struct orc_panel_3d {
    int id, level, parent_id, child1, child2, child3, child4, child5, child6, child7, child8;
    _Bool is_leaf;
    double max_x, min_x, max_y, min_y, min_z, max_z, radius;
    int point_count;
} orc_panel_3d;

/* interact_pair (structs.hpp:74-78); type 0=PP 1=PC 2=CP 3=CC. */
typedef struct orc_pair {
    int target_panel, source_panel, interact_type;
} orc_pair;

/* Tree produced by orc_tree_build. */
typedef struct orc_tree {
    int dim;
    int panel_count;
    int levels;
    int *level_start;       /* [levels+1] */
    orc_panel *panels;      /* [panel_count] */
    int *point_leaf_panel;  /* [N] */
    int *leaf_start;        /* [panel_count+1]: CSR offsets into leaf_points (leaves only non-empty) */
    int *leaf_points;       /* [N]: point indices of each leaf, in list order */
    int n_points;
} orc_tree;

/* Kernels. */
enum {
    ORC_KERNEL_POT = 0,       /* potential only, 1 component */
    ORC_KERNEL_REF_VEL3D = 1, /* reference "vel" kernel, 3D only, 3 components */
    ORC_KERNEL_POT_GRAD = 2   /* potential + target gradient, 1+dim components */
};

/* Options that the reference fixes implicitly. */
typedef struct orc_opts {
    int kernel;          /* ORC_KERNEL_* */
    int convention;      /* 0 = reference signs (poisson_fmm_impl.hpp:38,241); 1 = physical (negated) */
    int exclude_self;    /* 1 = skip pairs at zero distance (dx=dy=dz=0) */
    int deepest_only;    /* 1 = leaf_to_point only on the deepest level (downward_pass.cpp:14-16) */
    /* CPU-baseline sampling, >1: only every stride-th block of 64 consecutive
     * units is processed (list pairs; points, panels and leaves of the passes) */
    int near_stride;     /* PP/PC pairs */
    int far_stride;      /* CP/CC pairs */
    int pass_stride;     /* upward (orc_solve) and downward pass */
} orc_opts;

/* Seconds spent in the PP/PC loops of the last orc_fmm_interactions call, and
 * the fraction of PP/PC kernel evaluations that call performed; the same for
 * the CP/CC loops. */
double orc_last_near_seconds(void);
double orc_last_near_fraction(void);
double orc_last_far_seconds(void);
double orc_last_far_fraction(void);
/* Seconds and processed fraction of the last orc_solve's anterpolation,
 * child-to-parent, parent-to-child and leaf-to-point passes (4 each). */
void orc_last_passes(double *seconds, double *fraction);

int orc_components(int dim, int kernel);

/* ---- namelist (read_namelist.cpp:10-44) ---- */
int orc_read_run_config(const char *path, orc_run_config *cfg);

/* ---- basis (bli_impl.hpp:8-77) ---- */
void orc_interp_vals_bli(double *basis_vals, double x, double min_x, double max_x, int n);
void orc_bli_points_shift(double *pts, double min_x, double max_x, int n);

/* ---- tree (tree_construction.cpp:5-306) ---- */
int orc_tree_build(int dim, int thresh, int n, const double *x, const double *y, const double *z,
                   orc_tree *tree);
void orc_tree_free(orc_tree *tree);
void orc_pack_panels_2d(const orc_tree *tree, orc_panel_2d *out);
void orc_pack_panels_3d(const orc_tree *tree, orc_panel_3d *out);
/* padded [panel_count][thresh] leaf point lists, zero filled (tree_construction.cpp:286-295) */
void orc_panel_points_inside(const orc_tree *tree, int thresh, int *out);

/* ---- traversal (interaction_list.cpp:5-278) ---- */
/* Returns number of pairs; *list is malloc'd; counters written into cfg. */
int orc_dual_tree_traversal(orc_run_config *cfg, const orc_tree *tt, const orc_tree *ts, orc_pair **list);
void orc_split_interactions(const orc_pair *list, int count, orc_pair *pp, orc_pair *pc, orc_pair *cp,
                            orc_pair *cc);

/* ---- upward pass (upward_pass.cpp:7-29, upward_pass_impl.hpp:9-187) ---- */
void orc_upward_pass(const orc_run_config *cfg, const orc_tree *ts, const double *x, const double *y,
                     const double *z, const double *q, double *W /* [pc][P], accumulated */);

/* ---- interactions (poisson_fmm.cpp:8-26, poisson_vel_fmm.cpp:8-17) ---- */
void orc_fmm_interactions(const orc_run_config *cfg, const orc_opts *opt, const orc_tree *tt,
                          const orc_tree *ts, const double *xt, const double *yt, const double *zt,
                          const double *xs, const double *ys, const double *zs, const double *q,
                          const double *W, const orc_pair *list, int count,
                          double *out /* [C][Mt] accumulated */, double *Psi /* [C][pc_t][P] accumulated */);

/* ---- downward pass (downward_pass.cpp:7-41) ---- */
void orc_downward_pass(const orc_run_config *cfg, const orc_opts *opt, const orc_tree *tt,
                       const double *x, const double *y, const double *z, double *Psi /* [C][pc][P] */,
                       double *out /* [C][M] accumulated */);

/* ---- direct sum (direct_sum_impl.hpp:9-94) ---- */
/* targets: idx==NULL means all m targets, else the m indices in idx; out[C][m]. */
void orc_direct_sum(int dim, double eps, const orc_opts *opt, int m, const int *idx, const double *xt,
                    const double *yt, const double *zt, int n, const double *xs, const double *ys,
                    const double *zs, const double *q, double *out);

/* Whole pipeline (example_poisson_3d.cpp:70-160); phase seconds in t[5]
 * (tree, traversal, upward, interactions, downward). Returns 0 on success. */
int orc_solve(const orc_run_config *cfg_in, const orc_opts *opt, int m, const double *xt, const double *yt,
              const double *zt, int n, const double *xs, const double *ys, const double *zs,
              const double *q, int coincident, double *out, double *t, int *counts /* pp,pc,cp,cc */);

void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
