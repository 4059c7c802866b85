/*
 * btk.h -- C ABI of the MI355X barycentric Lagrange dual-tree treecode.
 *
 * One entry point per entry point of the reference's hot path, with the
 * reference's name prefixed by btk_.  Arguments are plain pointers and sizes:
 * device pointers (d_*) address HBM on the current HIP device, host pointers
 * are ordinary memory.  `stream` is a hipStream_t passed as void* (NULL is the
 * legacy default stream).  Every function returns 0 on success and a nonzero
 * status on failure; btk_last_error() then gives a thread-local message.  No
 * exception crosses this boundary.
 *
 * Records that the reference exchanges by value (RunConfig, TreePanel_2d,
 * TreePanel_3d, interact_pair, TreeInfo) have the reference's field order and
 * x86-64 layout, so arrays copied out of the library compare byte for byte
 * with the reference's own.
 *
 * Trees and interaction lists stay resident on the GPU behind opaque handles;
 * btk_tree_copy_* and btk_list_copy bring their reference-layout form to the
 * host on request.
 */
#ifndef BTK_H
#define BTK_H

#ifdef __cplusplus
extern "C" {
#endif

#define BTK_MAX_DEGREE 10 /* barytreek-config.h.in:20 */

/* One-byte boolean, as C++ bool in the reference's structs. */
#ifdef __cplusplus
#define BTK_BOOL bool
#else
#define BTK_BOOL _Bool
#endif

/* RunConfig (structs.hpp:6-21). */
typedef struct btk_run_config {
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
} btk_run_config;

/* TreeInfo (structs.hpp:23-27).  level_start is owned by the tree handle. */
typedef struct btk_tree_info {
    int panel_count;
    int levels;
    const int *level_start; /* [levels + 1] */
} btk_tree_info;

/* TreePanel_2d (structs.hpp:29-44), 80 bytes. */
typedef struct btk_panel_2d {
    int id, level, parent_id, child1, child2, child3, child4;
    BTK_BOOL is_leaf;
    double max_x, min_x, max_y, min_y, radius;
    int point_count;
} btk_panel_2d;

/* TreePanel_3d (structs.hpp:46-67), 112 bytes. */
typedef struct btk_panel_3d {
    int id, level, parent_id, child1, child2, child3, child4, child5, child6, child7, child8;
    BTK_BOOL is_leaf;
    double max_x, min_x, max_y, min_y, min_z, max_z, radius;
    int point_count;
} btk_panel_3d;

/* interact_pair (structs.hpp:74-78); type 0 = PP, 1 = PC, 2 = CP, 3 = CC. */
typedef struct btk_interact_pair {
    int target_panel;
    int source_panel;
    int interact_type;
} btk_interact_pair;

typedef struct btk_tree btk_tree; /* device-resident tree */
typedef struct btk_list btk_list; /* device-resident interaction list */

/* Interaction kernels.  Signs follow the reference unless convention = 1. */
enum {
    BTK_KERNEL_POT = 0,       /* 2D (1/4pi) log(r^2+eps^2); 3D -1/(4pi sqrt(r^2+eps^2)) (poisson_fmm_impl.hpp:38,241) */
    BTK_KERNEL_REF_VEL3D = 1, /* 3D q (-1/4pi) (t-s)/(r^2+eps^2), 3 outputs (poisson_vel_fmm_impl.hpp:42,53) */
    BTK_KERNEL_POT_GRAD = 2   /* potential and its target gradient, 1 + dim outputs (no reference counterpart) */
};

enum {
    BTK_PREC_FP64 = 0,  /* every evaluation in FP64 */
    BTK_PREC_MIXED = 1  /* FP32 kernel evaluations, FP64 geometry, weights and per-target sums */
};

enum {
    BTK_SUM_BLOCKED = 0, /* FP32 partial sums over blocks of 64 sources, each added into FP64 per target */
    BTK_SUM_KAHAN = 1    /* the same blocks, each summed with Kahan's compensation (one target per lane) */
};

enum {
    BTK_SCHEDULE_GROUPS = 0, /* one launch per group of work items (near or far, one degree) */
    BTK_SCHEDULE_QUEUE = 1   /* one persistent launch drains every group's work items from one queue, biggest
                                first, by atomic ticket (no reference counterpart; the reference launches one
                                parallel_for per interaction type, poisson_fmm.cpp:21-25) */
};

/* Choices the reference makes implicitly.  Zero-initialised = reference behaviour. */
typedef struct btk_options {
    int kernel;       /* BTK_KERNEL_* */
    int convention;   /* 0 = reference signs; 1 = physical signs (both kernels negated) */
    int exclude_self; /* 1 = skip pairs at zero distance; 0 = include them (reference) */
    int deepest_only; /* 1 = leaf_to_point on the deepest level only (downward_pass.cpp:26-28) */
    int precision;    /* BTK_PREC_* */
    int mac;          /* 0 = the reference's fixed theta MAC; 1 = the learned MAC (btk_mac_*) where it accepts */
    int schedule;     /* BTK_SCHEDULE_* of the PP/PC/CP/CC work items */
    int summation;    /* BTK_SUM_*: how mixed-precision PP/PC sums are kept in FP32 */
} btk_options;

/* Number of output components of a kernel: 1, 3, or 1 + dim. */
int btk_components(int dim, int kernel);

/* Thread-local message of the last failure on this thread. */
const char *btk_last_error(void);

/* Releases the device memory the library keeps cached for reuse. */
int btk_release_cached(void);

/* Launch shapes of the interaction kernels (no reference counterpart; the
 * reference has one thread per pair).  Process-wide; zero fields mean the
 * built-in defaults.  In FP64, results depend on the shapes only through the
 * order of FP64 atomic adds; in mixed precision the grouping of FP32 partial
 * sums changes as well (differences at FP32 rounding level). */
typedef struct btk_tuning {
    int near_items; /* work items the PP+PC launch aims for (entry slices are added to reach it) */
    int far_items;  /* the same for the CP+CC launch */
    int near_tpt;   /* targets per lane of the mixed-precision PP+PC kernel: 1 or 2 */
    int far_tpt;    /* the same for CP+CC */
    int mutual;     /* 1 = mixed precision, one tree: a PP leaf pair listed in both orders, (T,S) and (S,T),
                       is evaluated once for both directions; 0 = once per entry (default) */
    int cc_group;   /* mixed precision, 3D, potential (+gradient): CC pairs with the same two levels and
                       box-centre offset share their kernel matrix; a set of at least cc_group such pairs
                       runs as one matrix-core product (split-fp16 operands, FP32 accuracy per product;
                       each target panel's pairs summed in FP64 in a fixed order).  0 = default (4),
                       -1 = off */
    int near_waves; /* wavefronts per workgroup of the PP+PC kernel, 1, 2 or 4 (0 = default: 4 in 3D, 1 in
                       2D): they take
                       consecutive blocks of a panel's targets and share one LDS tile of 64 * near_waves
                       sources, so each source is read from memory once per workgroup */
    int far_waves;  /* the same for CP+CC (0 = default, 1) */
    int multi_vectors; /* btk_solve_multi's MFMA kernels: charge vectors per workgroup, 32, 64 or 128
                          (0 = default, 64); a workgroup of multi_vectors / 16 wavefronts takes 64 targets
                          and evaluates each kernel value once per 32 vectors */
} btk_tuning;
int btk_set_tuning(const btk_tuning *t);
int btk_get_tuning(btk_tuning *t);

/* Autotuner (no reference counterpart): builds the trees and list of one
 * problem (the arguments of btk_solve), times the PP+PC and CP+CC launches
 * under each candidate shape (work-item targets 8192..131072, 1, 2 or 4
 * wavefronts per workgroup; in mixed precision also 1 or 2 targets per lane
 * and mutual on or off), installs
 * the fastest with btk_set_tuning and returns it in *chosen (may be null).
 * Writes no caller memory besides *chosen. */
int btk_autotune(const btk_run_config *cfg, const btk_options *opts, int m, const double *d_xt, const double *d_yt,
                 const double *d_zt, int n, const double *d_xs, const double *d_ys, const double *d_zs,
                 const double *d_q, int coincident, btk_tuning *chosen, void *stream);

/* Autotuner of btk_solve_multi (no reference counterpart): with the arguments
 * of btk_solve_multi (d_q holds nrhs charge vectors of n sources), builds the
 * trees, list and proxy weights once, times the many-vector PP+PC and CP+CC
 * launches with 32, 64 and 128 charge vectors per workgroup, installs the
 * fastest (btk_tuning.multi_vectors; the other fields are kept) and returns
 * the tuning in *chosen (may be null).  Needs mixed precision and nrhs >= 16,
 * the case btk_solve_multi runs on MFMA.  Writes no caller memory besides
 * *chosen. */
int btk_autotune_multi(const btk_run_config *cfg, const btk_options *opts, int m, const double *d_xt,
                       const double *d_yt, const double *d_zt, int n, const double *d_xs, const double *d_ys,
                       const double *d_zs, const double *d_q, int nrhs, int coincident, btk_tuning *chosen,
                       void *stream);

/* read_run_config (read_namelist.hpp:7, read_namelist.cpp:10-44).
 * cfg->dim must be set before the call, as in the reference drivers. */
int btk_read_run_config(const char *path, btk_run_config *cfg);

/* blfmm_tree_construction_2d/_3d (tree_construction.hpp:7-9).  Builds the
 * tree on the GPU from n device-resident coordinates (original order). */
int btk_tree_construction_2d(const btk_run_config *cfg, const double *d_x, const double *d_y, int n,
                             void *stream, btk_tree **tree);
int btk_tree_construction_3d(const btk_run_config *cfg, const double *d_x, const double *d_y,
                             const double *d_z, int n, void *stream, btk_tree **tree);
int btk_tree_get_info(const btk_tree *tree, btk_tree_info *info);
/* blfmm_panels in the reference's AoS layout (host array of panel_count). */
int btk_tree_copy_panels_2d(const btk_tree *tree, btk_panel_2d *panels);
int btk_tree_copy_panels_3d(const btk_tree *tree, btk_panel_3d *panels);
/* point_leaf_panel (host int[n]). */
int btk_tree_copy_point_leaf_panel(const btk_tree *tree, int *point_leaf_panel);
/* panel_points_inside, padded [panel_count][width], zero filled, leaf rows only
 * (tree_construction.cpp:286-295); width is the reference's cluster_thresh. */
int btk_tree_copy_panel_points_inside(const btk_tree *tree, int width, int *panel_points_inside);
/* The same lists in compact form: perm[n] (point indices, leaf by leaf,
 * ascending within a leaf) and start[panel_count + 1]; every panel, leaf or
 * not, owns perm[start[p] .. start[p] + point_count). */
int btk_tree_copy_permutation(const btk_tree *tree, int *perm, int *start);
void btk_tree_destroy(btk_tree *tree);

/* dual_tree_traversal_2d/_3d (interaction_list.hpp:6-8).  Writes the five
 * counters into cfg as the reference does (interaction_list.cpp:242-249). */
int btk_dual_tree_traversal_2d(btk_run_config *cfg, const btk_tree *target, const btk_tree *source, void *stream,
                               btk_list **list);
int btk_dual_tree_traversal_3d(btk_run_config *cfg, const btk_tree *target, const btk_tree *source, void *stream,
                               btk_list **list);
/* The 3D traversal with the installed learned MAC (btk_mac_train /
 * btk_mac_set_weights and btk_mac_set_budget; no counterpart in the
 * reference; the decision rule is described with btk_mac_sample below):
 * each pair's type and degree among 4, 6, 8, 10 (up to cfg->interp_degree,
 * which must be one of them).  interact_type holds the type in bits 0-1 and
 * the degree index (degree 4 + 2 * index) in bits 2-3.  Such a
 * list is for inspection and btk_split_interactions; btk_solve with
 * btk_options.mac runs it.  Counters as btk_dual_tree_traversal_3d. */
int btk_dual_tree_traversal_learned_3d(btk_run_config *cfg, const btk_tree *target, const btk_tree *source,
                                       void *stream, btk_list **list);
int btk_list_count(const btk_list *list);
/* The full list in the reference's FIFO order (host array of count). */
int btk_list_copy(const btk_list *list, btk_interact_pair *pairs);
/* split_interactions (interaction_list.hpp:10): stable split by type
 * (interact_type & 3). */
int btk_split_interactions(const btk_list *list, btk_interact_pair *pp, btk_interact_pair *pc,
                           btk_interact_pair *cp, btk_interact_pair *cc);
void btk_list_destroy(btk_list *list);

/* upward_pass_2d/_3d (upward_pass.hpp:7-9).  d_charges in original point
 * order; writes d_proxy_source_weights[panel_count][P], P = (n+1)^dim, for
 * every panel (no need to zero it first). */
int btk_upward_pass_2d(const btk_run_config *cfg, const btk_tree *source, const double *d_charges,
                       double *d_proxy_source_weights, void *stream);
int btk_upward_pass_3d(const btk_run_config *cfg, const btk_tree *source, const double *d_charges,
                       double *d_proxy_source_weights, void *stream);

/* No counterpart in the reference: btk_upward_pass_3d for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or
 * allocating library routine: the charges in the tree's sorted order, the one
 * scratch buffer of the upward pass, live in a fixed device array of
 * BTK_NOALLOC_MAX_POINTS doubles (64 MiB) that each GPU the library runs on
 * holds until the process ends.  A source tree of more points is refused
 * with an error.  Same result as btk_upward_pass_3d.  Unlike it, the call
 * does not synchronize: the work stays queued on `stream`.  Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (on different
 * streams, or before an earlier call's work on another stream has finished)
 * race; calls on one stream are ordered. */
#define BTK_NOALLOC_MAX_POINTS 8388608
int btk_upward_pass_3d_noalloc(const btk_run_config *cfg, const btk_tree *source, const double *d_charges,
                               double *d_proxy_source_weights, void *stream);

/* poisson_fmm_interactions_2d/_3d (poisson_fmm.hpp:9-15) and
 * poisson_vel_fmm_interactions_3d (poisson_vel_fmm.hpp:9-11), the kernel
 * chosen by opts.  Adds PP and PC to d_soln[C][M] (original target order)
 * and CP and CC to d_proxy_target_weights[C][panel_count_target][P]. */
int btk_fmm_interactions_2d(const btk_run_config *cfg, const btk_options *opts, const btk_tree *target,
                            const btk_tree *source, const btk_list *list, const double *d_charges,
                            const double *d_proxy_source_weights, double *d_soln,
                            double *d_proxy_target_weights, void *stream);
int btk_fmm_interactions_3d(const btk_run_config *cfg, const btk_options *opts, const btk_tree *target,
                            const btk_tree *source, const btk_list *list, const double *d_charges,
                            const double *d_proxy_source_weights, double *d_soln,
                            double *d_proxy_target_weights, void *stream);

/* downward_pass_2d/_3d/_3d_3 (downward_pass.hpp:7-11): parent_to_child
 * through the levels, then leaf_to_point adding into d_soln[C][M]. */
int btk_downward_pass_2d(const btk_run_config *cfg, const btk_options *opts, const btk_tree *target,
                         double *d_proxy_target_weights, double *d_soln, void *stream);
int btk_downward_pass_3d(const btk_run_config *cfg, const btk_options *opts, const btk_tree *target,
                         double *d_proxy_target_weights, double *d_soln, void *stream);

/* poisson_2d_dir_sum / poisson_3d_dir_sum / poisson_vel_3d_dir_sum
 * (direct_sum_impl.hpp:9-94), FP64.  Adds into d_out[C][m]; target a is point
 * d_target_idx[a] (or a when d_target_idx is NULL). */
int btk_direct_sum_2d(const btk_options *opts, double eps, int m, const int *d_target_idx, const double *d_xt,
                      const double *d_yt, int n, const double *d_xs, const double *d_ys, const double *d_q,
                      double *d_out, void *stream);
int btk_direct_sum_3d(const btk_options *opts, double eps, int m, const int *d_target_idx, const double *d_xt,
                      const double *d_yt, const double *d_zt, int n, const double *d_xs, const double *d_ys,
                      const double *d_zs, const double *d_q, double *d_out, void *stream);

/* Phase times (seconds, from HIP events) and counts of one btk_solve. */
typedef struct btk_solve_stats {
    double t_tree, t_traversal, t_upward, t_interact, t_downward, t_total;
    long long count[4];  /* PP, PC, CP, CC pairs */
    double evals[4];     /* kernel evaluations per type (sum n_t n_s, n_t P, P n_s, P^2) */
    double t_kernel_near; /* device time of the PP+PC kernel (HIP events on the solve's stream) */
    double t_kernel_far;  /* device time of the CP+CC kernel */
    int panel_count_target, panel_count_source, levels_target, levels_source;
    long long cc_shared; /* CC pairs evaluated as shared-kernel-matrix products (btk_tuning.cc_group) */
} btk_solve_stats;

/* The driver's sequence (example_poisson_3d.cpp:70-160) as one call: tree(s),
 * traversal, upward, interactions, downward.  Adds into d_out[C][m].
 * coincident != 0 means targets and sources are the same points (one tree). */
int btk_solve(const btk_run_config *cfg, const btk_options *opts, int m, const double *d_xt, const double *d_yt,
              const double *d_zt, int n, const double *d_xs, const double *d_ys, const double *d_zs,
              const double *d_q, int coincident, double *d_out, btk_solve_stats *stats, void *stream);

/* Many charge vectors over one geometry (no reference counterpart): btk_solve
 * for nrhs vectors d_q[nrhs][n] sharing one tree, list and set of kernel
 * evaluations.  Adds into d_out[nrhs][C][m].  With BTK_PREC_MIXED and
 * nrhs >= 16 the interactions run as FP32 MFMA products (kernel tile x charge
 * tile); otherwise each vector runs the single-vector interaction kernels.
 * Kernels: BTK_KERNEL_POT or BTK_KERNEL_POT_GRAD. */
int btk_solve_multi(const btk_run_config *cfg, const btk_options *opts, int m, const double *d_xt, const double *d_yt,
                    const double *d_zt, int n, const double *d_xs, const double *d_ys, const double *d_zs,
                    const double *d_q, int nrhs, int coincident, double *d_out, btk_solve_stats *stats,
                    void *stream);

/* Learned MAC (no reference counterpart).  A network of 10 -> 64 -> 64 -> 12
 * (ReLU; FP32; BTK_MAC_WEIGHTS parameters: W1[64][10], b1[64], W2[64][64],
 * b2[64], W3[12][64], b3[12]) reads 10 features of a candidate cluster pair
 * (separation ratio (r_t + r_s)/d, r_t/d, r_s/d, log2 of both point counts,
 * both levels, log2(r_t/r_s), g/d and (r_t + r_s)/g with g the distance
 * between the boxes) and predicts log10 of the relative error (L2 over the
 * potential and its gradient) of its PC, CP and CC approximations at degrees
 * 4, 6, 8 and 10 (output 4*type' + degree
 * index, type' = PC 0, CP 1, CC 2; scaled by 1/4, clamped to [-14, 0]).
 * With opts->mac = 1 (3D only; interp_degree 4, 6, 8 or 10, the highest
 * degree used) the traversal evaluates it on the non-PP pairs whose boxes are
 * apart (a positive gap) with separation ratio in [0.3, 1.2), the network's
 * domain.  Its candidates for a pair are the approximations at degrees up to
 * interp_degree whose predicted error is within the budget: CC for any two
 * panels, PC and CP where the panel on the point side is a leaf (at most
 * fmm_cluster_thresh points, as in the reference's types).  A pair the theta
 * MAC accepts takes the cheapest (fewest kernel evaluations) of these and the
 * theta MAC's own choice (its count-given type at interp_degree).  A pair the
 * theta MAC would refine takes the cheapest candidate that costs fewer
 * evaluations than the pair's direct sum; otherwise the theta MAC refines
 * it, as it decides every pair outside the domain.  The solve runs
 * the upward pass at interp_degree, restricts the charges to each lower
 * degree, and extends each lower degree's potentials into interp_degree's
 * before one downward pass (exact polynomial identities, to rounding).  The
 * model and budget are process-wide. */
#define BTK_MAC_FEATURES 10
#define BTK_MAC_OUTPUTS 12
#define BTK_MAC_WEIGHTS 5644
/* Training samples from the treecode's own traversal of the given 3D points:
 * n_samples pairs drawn from the pairs the traversal visits whose boxes are
 * apart with separation ratio in [0.3, 1.2), each with its features
 * d_X[n][10] and the measured errors
 * d_Y[n][12] of its approximations (relative L2 over the potential 1/r and
 * the gradient kernel d/r^3, each component interpolated as the treecode
 * does) against the direct sum over up to 64 target and 1024 source points
 * of the pair (charges U[-1,1] from seed).
 * d_pairs[n][2] (may be null) receives each sample's (target, source) panel
 * ids.  *count receives the number of samples written (<= n_samples). */
int btk_mac_sample(const btk_run_config *cfg, int n, const double *d_x, const double *d_y, const double *d_z,
                   int n_samples, unsigned seed, float *d_X, float *d_Y, int *d_pairs, int *count, void *stream);
/* Adam on the mean squared error of the scaled outputs, minibatches of
 * `batch`, from the installed weights (random He initialisation from seed
 * when none are installed or reinit != 0); installs the result; *loss gets
 * the last epoch's mean loss. */
int btk_mac_train(const float *d_X, const float *d_Y, int n, int epochs, int batch, float lr, unsigned seed,
                  int reinit, float *loss, void *stream);
int btk_mac_set_weights(const float *h_weights, int count);
int btk_mac_get_weights(float *h_weights, int count);
/* Maximum predicted relative error of an accepted pair (default 1e-7). */
int btk_mac_set_budget(double rel_err);
/* Relative cost of one kernel evaluation of a PP, PC, CP and CC pair
 * (costs[0..3], each > 0; default all 1, so that the learned MAC minimises
 * evaluations).  The learned MAC takes, among the candidates within the
 * budget, the one of least weighted cost: with CC products on the matrix
 * cores (mixed precision, btk_tuning.cc_group) a CC evaluation costs about
 * an eighth of a near-field one on the vector ALU (MI355X, 3D potential and
 * gradient: 0.048 against 0.40 ns).  Process-wide, like the budget. */
int btk_mac_set_costs(const float *costs);
/* The installed network on d_X[n][10]: d_out[n][12] (scaled outputs). */
int btk_mac_forward(const float *d_X, int n, float *d_out, void *stream);
/* Loss and its gradient (BTK_MAC_WEIGHTS values, the weights' layout) for
 * the installed weights on one batch. */
int btk_mac_gradients(const float *d_X, const float *d_Y, int n, float *d_grad, float *loss, void *stream);

/* A solve plan for one geometry (no reference counterpart; for iterative
 * methods that apply one operator to many vectors): the trees, the list and
 * the interaction work items are built once, with the tuning current at
 * creation.  btk_plan_execute then runs gather, upward pass, interactions
 * and downward pass for one charge vector d_q[n] as a single HIP graph launch
 * on `stream`, adding into d_out[C][m]; the graph is captured (on a stream
 * the plan owns) at the first call and again whenever d_q or d_out differs
 * from the previous call.  Calls on one plan must not overlap. */
typedef struct btk_plan btk_plan;
int btk_plan_create(const btk_run_config *cfg, const btk_options *opts, int m, const double *d_xt, const double *d_yt,
                    const double *d_zt, int n, const double *d_xs, const double *d_ys, const double *d_zs,
                    int coincident, void *stream, btk_plan **plan);
int btk_plan_execute(btk_plan *plan, const double *d_q, double *d_out, void *stream);
void btk_plan_destroy(btk_plan *plan);

/* Many independent problems in one solve (no reference counterpart): problem
 * p owns points [offsets[p], offsets[p+1]) of d_x/d_y/d_z/d_q (targets =
 * sources within a problem; offsets is a host array of nprob + 1 entries,
 * offsets[0] = 0, every problem non-empty).  The problems form one forest --
 * each problem's tree, list and result are those of its own btk_solve -- and
 * every phase runs as one launch over all problems.  Adds into
 * d_out[C][offsets[nprob]].  deepest_only is rejected when nprob > 1. */
int btk_solve_batch(const btk_run_config *cfg, const btk_options *opts, int nprob, const int *offsets,
                    const double *d_x, const double *d_y, const double *d_z, const double *d_q, double *d_out,
                    btk_solve_stats *stats, void *stream);

/* A solve plan for a fixed forest of nprob problems (no reference
 * counterpart): btk_solve_batch's forest, list and work items built once, then
 * btk_plan_execute replays gather, upward pass, interactions and downward
 * pass as one HIP graph, adding into d_out[C][offsets[nprob]].  With
 * opts->schedule = BTK_SCHEDULE_QUEUE the interactions of every problem are
 * one persistent launch. */
int btk_plan_create_batch(const btk_run_config *cfg, const btk_options *opts, int nprob, const int *offsets,
                          const double *d_x, const double *d_y, const double *d_z, void *stream, btk_plan **plan);

#ifdef __cplusplus
}
#endif
#endif
