"""Python binding of libbtk, the MI355X barycentric Lagrange dual-tree treecode.

The product is the C ABI in include/btk.h (libbtk.so, hand-written HIP
kernels).  This module binds it with ctypes for tests, benchmarks and Python
callers; device memory comes from torch tensors on the GPU.  Loading fails
loudly when libbtk.so has not been built -- there is no fallback path.

Function names mirror the reference's entry points (read_run_config,
blfmm_tree_construction_3d, dual_tree_traversal_3d, split_interactions,
upward_pass_3d, poisson_fmm_interactions_3d, poisson_vel_fmm_interactions_3d,
downward_pass_3d, downward_pass_3d_3 and their 2D forms).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbtk.so")

KERNEL_POT, KERNEL_REF_VEL3D, KERNEL_POT_GRAD = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
SCHEDULE_GROUPS, SCHEDULE_QUEUE = 0, 1
SUM_BLOCKED, SUM_KAHAN = 0, 1

EXPORTED = [
    "btk_components", "btk_last_error", "btk_release_cached", "btk_set_tuning", "btk_get_tuning",
    "btk_autotune", "btk_autotune_multi",
    "btk_read_run_config",
    "btk_tree_construction_2d", "btk_tree_construction_3d", "btk_tree_get_info",
    "btk_tree_copy_panels_2d", "btk_tree_copy_panels_3d", "btk_tree_copy_point_leaf_panel",
    "btk_tree_copy_panel_points_inside", "btk_tree_copy_permutation", "btk_tree_destroy",
    "btk_dual_tree_traversal_2d", "btk_dual_tree_traversal_3d", "btk_dual_tree_traversal_learned_3d",
    "btk_list_count", "btk_list_copy",
    "btk_split_interactions", "btk_list_destroy", "btk_upward_pass_2d", "btk_upward_pass_3d",
    "btk_upward_pass_3d_noalloc",
    "btk_fmm_interactions_2d", "btk_fmm_interactions_3d", "btk_downward_pass_2d", "btk_downward_pass_3d",
    "btk_direct_sum_2d", "btk_direct_sum_3d", "btk_solve", "btk_solve_multi", "btk_solve_batch",
    "btk_plan_create", "btk_plan_create_batch", "btk_plan_execute", "btk_plan_destroy",
    "btk_mac_sample", "btk_mac_train", "btk_mac_set_weights", "btk_mac_get_weights", "btk_mac_set_budget",
    "btk_mac_set_costs",
    "btk_mac_forward", "btk_mac_gradients",
]


class BtkError(RuntimeError):
    pass


class RunConfig(C.Structure):
    """btk_run_config == RunConfig (structs.hpp:6-21)."""
    _fields_ = [
        ("dim", C.c_int), ("interp_degree", C.c_int), ("interp_point_count", C.c_int),
        ("fmm_theta", C.c_double), ("fmm_cluster_thresh", C.c_int),
        ("fmm_interaction_count", C.c_int), ("fmm_pp_count", C.c_int), ("fmm_pc_count", C.c_int),
        ("fmm_cp_count", C.c_int), ("fmm_cc_count", C.c_int), ("ker_eps", C.c_double),
        ("mpi_p", C.c_int), ("mpi_id", C.c_int),
    ]


class TreeInfo(C.Structure):
    _fields_ = [("panel_count", C.c_int), ("levels", C.c_int), ("level_start", C.POINTER(C.c_int))]


class Options(C.Structure):
    _fields_ = [("kernel", C.c_int), ("convention", C.c_int), ("exclude_self", C.c_int),
                ("deepest_only", C.c_int), ("precision", C.c_int), ("mac", C.c_int), ("schedule", C.c_int),
                ("summation", C.c_int)]


class Tuning(C.Structure):
    """btk_tuning: launch shapes of the interaction kernels (0 = default)."""
    _fields_ = [("near_items", C.c_int), ("far_items", C.c_int), ("near_tpt", C.c_int), ("far_tpt", C.c_int),
                ("mutual", C.c_int), ("cc_group", C.c_int), ("near_waves", C.c_int), ("far_waves", C.c_int),
                ("multi_vectors", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("t_tree", C.c_double), ("t_traversal", C.c_double), ("t_upward", C.c_double),
                ("t_interact", C.c_double), ("t_downward", C.c_double), ("t_total", C.c_double),
                ("count", C.c_longlong * 4), ("evals", C.c_double * 4), ("t_kernel_near", C.c_double),
                ("t_kernel_far", C.c_double), ("panel_count_target", C.c_int),
                ("panel_count_source", C.c_int), ("levels_target", C.c_int), ("levels_source", C.c_int),
                ("cc_shared", C.c_longlong)]


PAIR_DTYPE = np.dtype([("target_panel", "<i4"), ("source_panel", "<i4"), ("interact_type", "<i4")])
PANEL3D_DTYPE = np.dtype({
    "names": ["id", "level", "parent_id", "child1", "child2", "child3", "child4", "child5", "child6",
              "child7", "child8", "is_leaf", "max_x", "min_x", "max_y", "min_y", "min_z", "max_z",
              "radius", "point_count"],
    "formats": ["<i4"] * 11 + ["u1"] + ["<f8"] * 7 + ["<i4"],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44, 48, 56, 64, 72, 80, 88, 96, 104],
    "itemsize": 112,
})
PANEL2D_DTYPE = np.dtype({
    "names": ["id", "level", "parent_id", "child1", "child2", "child3", "child4", "is_leaf", "max_x",
              "min_x", "max_y", "min_y", "radius", "point_count"],
    "formats": ["<i4"] * 7 + ["u1"] + ["<f8"] * 5 + ["<i4"],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64, 72],
    "itemsize": 80,
})

_lib = None


def lib():
    """Loads libbtk.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BtkError(f"{LIB_PATH} is not built; run `python -m barytreek_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.c_void_p
    cfgp = C.POINTER(RunConfig)
    optp = C.POINTER(Options)
    L.btk_components.argtypes = [C.c_int, C.c_int]
    L.btk_last_error.restype = C.c_char_p
    L.btk_read_run_config.argtypes = [C.c_char_p, cfgp]
    L.btk_set_tuning.argtypes = [C.POINTER(Tuning)]
    L.btk_get_tuning.argtypes = [C.POINTER(Tuning)]
    L.btk_autotune.argtypes = [cfgp, optp, C.c_int, dp, dp, dp, C.c_int, dp, dp, dp, dp, C.c_int,
                               C.POINTER(Tuning), vp]
    L.btk_autotune_multi.argtypes = [cfgp, optp, C.c_int, dp, dp, dp, C.c_int, dp, dp, dp, dp, C.c_int, C.c_int,
                                     C.POINTER(Tuning), vp]
    L.btk_tree_construction_2d.argtypes = [cfgp, dp, dp, C.c_int, vp, C.POINTER(vp)]
    L.btk_tree_construction_3d.argtypes = [cfgp, dp, dp, dp, C.c_int, vp, C.POINTER(vp)]
    L.btk_tree_get_info.argtypes = [vp, C.POINTER(TreeInfo)]
    L.btk_tree_copy_panels_2d.argtypes = [vp, vp]
    L.btk_tree_copy_panels_3d.argtypes = [vp, vp]
    L.btk_tree_copy_point_leaf_panel.argtypes = [vp, vp]
    L.btk_tree_copy_panel_points_inside.argtypes = [vp, C.c_int, vp]
    L.btk_tree_copy_permutation.argtypes = [vp, vp, vp]
    L.btk_tree_destroy.argtypes = [vp]
    L.btk_tree_destroy.restype = None
    L.btk_dual_tree_traversal_2d.argtypes = [cfgp, vp, vp, vp, C.POINTER(vp)]
    L.btk_dual_tree_traversal_3d.argtypes = [cfgp, vp, vp, vp, C.POINTER(vp)]
    L.btk_dual_tree_traversal_learned_3d.argtypes = [cfgp, vp, vp, vp, C.POINTER(vp)]
    L.btk_list_count.argtypes = [vp]
    L.btk_list_copy.argtypes = [vp, vp]
    L.btk_split_interactions.argtypes = [vp, vp, vp, vp, vp]
    L.btk_list_destroy.argtypes = [vp]
    L.btk_list_destroy.restype = None
    L.btk_upward_pass_2d.argtypes = [cfgp, vp, dp, dp, vp]
    L.btk_upward_pass_3d.argtypes = [cfgp, vp, dp, dp, vp]
    L.btk_upward_pass_3d_noalloc.argtypes = [cfgp, vp, dp, dp, vp]
    L.btk_fmm_interactions_2d.argtypes = [cfgp, optp, vp, vp, vp, dp, dp, dp, dp, vp]
    L.btk_fmm_interactions_3d.argtypes = [cfgp, optp, vp, vp, vp, dp, dp, dp, dp, vp]
    L.btk_downward_pass_2d.argtypes = [cfgp, optp, vp, dp, dp, vp]
    L.btk_downward_pass_3d.argtypes = [cfgp, optp, vp, dp, dp, vp]
    L.btk_direct_sum_2d.argtypes = [optp, C.c_double, C.c_int, dp, dp, dp, C.c_int, dp, dp, dp, dp, vp]
    L.btk_direct_sum_3d.argtypes = [optp, C.c_double, C.c_int, dp, dp, dp, dp, C.c_int, dp, dp, dp, dp, dp, vp]
    L.btk_solve.argtypes = [cfgp, optp, C.c_int, dp, dp, dp, C.c_int, dp, dp, dp, dp, C.c_int, dp,
                            C.POINTER(SolveStats), vp]
    L.btk_solve_multi.argtypes = [cfgp, optp, C.c_int, dp, dp, dp, C.c_int, dp, dp, dp, dp, C.c_int, C.c_int, dp,
                                  C.POINTER(SolveStats), vp]
    L.btk_solve_batch.argtypes = [cfgp, optp, C.c_int, ip, dp, dp, dp, dp, dp, C.POINTER(SolveStats), vp]
    fp = C.c_void_p
    L.btk_mac_sample.argtypes = [cfgp, C.c_int, dp, dp, dp, C.c_int, C.c_uint, fp, fp, vp, C.POINTER(C.c_int), vp]
    L.btk_mac_train.argtypes = [fp, fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint, C.c_int,
                                C.POINTER(C.c_float), vp]
    L.btk_mac_set_weights.argtypes = [C.POINTER(C.c_float), C.c_int]
    L.btk_mac_get_weights.argtypes = [C.POINTER(C.c_float), C.c_int]
    L.btk_mac_set_budget.argtypes = [C.c_double]
    L.btk_mac_set_costs.argtypes = [C.POINTER(C.c_float)]
    L.btk_mac_forward.argtypes = [fp, C.c_int, fp, vp]
    L.btk_mac_gradients.argtypes = [fp, fp, C.c_int, fp, C.POINTER(C.c_float), vp]
    L.btk_plan_create.argtypes = [cfgp, optp, C.c_int, dp, dp, dp, C.c_int, dp, dp, dp, C.c_int, vp, C.POINTER(vp)]
    L.btk_plan_create_batch.argtypes = [cfgp, optp, C.c_int, ip, dp, dp, dp, vp, C.POINTER(vp)]
    L.btk_plan_execute.argtypes = [vp, dp, dp, vp]
    L.btk_plan_destroy.argtypes = [vp]
    L.btk_plan_destroy.restype = None
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise BtkError(lib().btk_last_error().decode())


def components(dim: int, kernel: int) -> int:
    return lib().btk_components(dim, kernel)


def get_tuning() -> Tuning:
    t = Tuning()
    _check(lib().btk_get_tuning(C.byref(t)))
    return t


def set_tuning(near_items=0, far_items=0, near_tpt=0, far_tpt=0, mutual=0, cc_group=0, near_waves=0,
               far_waves=0, multi_vectors=0) -> None:
    """Process-wide launch shapes of the interaction kernels; 0 = default
    (cc_group: -1 turns the shared-matrix CC products off; multi_vectors:
    charge vectors per workgroup of the many-vector MFMA kernels)."""
    t = Tuning()
    t.near_items, t.far_items, t.near_tpt, t.far_tpt = near_items, far_items, near_tpt, far_tpt
    t.mutual, t.cc_group, t.near_waves, t.far_waves = mutual, cc_group, near_waves, far_waves
    t.multi_vectors = multi_vectors
    _check(lib().btk_set_tuning(C.byref(t)))


def restore_tuning(t: Tuning) -> None:
    """Installs a record from get_tuning() again."""
    _check(lib().btk_set_tuning(C.byref(t)))


def autotune(cfg, opts, xt, yt, zt, q, xs=None, ys=None, zs=None, stream=None) -> Tuning:
    """Times the interaction launches of this problem under candidate shapes
    and installs the fastest process-wide (btk_autotune)."""
    coincident = xs is None
    if coincident:
        xs, ys, zs = xt, yt, zt
    t = Tuning()
    _check(lib().btk_autotune(C.byref(cfg), C.byref(opts), int(xt.numel()), _ptr(xt), _ptr(yt), _ptr(zt),
                              int(xs.numel()), _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q), int(coincident), C.byref(t),
                              _stream(stream)))
    return t


def make_config(dim, degree, thresh, theta=0.7, eps=-1.0) -> RunConfig:
    c = RunConfig()
    c.dim, c.interp_degree = dim, degree
    c.interp_point_count = (degree + 1) ** dim
    c.fmm_theta, c.fmm_cluster_thresh, c.ker_eps = theta, thresh, eps
    c.mpi_p, c.mpi_id = 1, 0
    return c


def options(kernel=KERNEL_POT, convention=0, exclude_self=0, deepest_only=0, precision=PREC_FP64, mac=0,
            schedule=0, summation=0) -> Options:
    o = Options()
    o.kernel, o.convention, o.exclude_self = kernel, convention, exclude_self
    o.deepest_only, o.precision, o.mac, o.schedule = deepest_only, precision, mac, schedule
    o.summation = summation
    return o


def read_run_config(path: str, dim: int) -> RunConfig:
    """read_run_config (read_namelist.cpp:10-44); defaults of structs.hpp:10,17."""
    c = RunConfig()
    c.dim, c.fmm_theta, c.ker_eps = dim, 0.7, -1.0
    _check(lib().btk_read_run_config(path.encode(), C.byref(c)))
    return c


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Tree:
    """A device-resident tree (btk_tree handle)."""

    def __init__(self, cfg: RunConfig, x, y, z=None, stream=None):
        self.dim = cfg.dim
        self.n = int(x.numel())
        h = C.c_void_p()
        if cfg.dim == 3:
            _check(lib().btk_tree_construction_3d(C.byref(cfg), _ptr(x), _ptr(y), _ptr(z), self.n,
                                                  _stream(stream), C.byref(h)))
        else:
            _check(lib().btk_tree_construction_2d(C.byref(cfg), _ptr(x), _ptr(y), self.n, _stream(stream),
                                                  C.byref(h)))
        self.h = h
        info = TreeInfo()
        _check(lib().btk_tree_get_info(self.h, C.byref(info)))
        self.panel_count, self.levels = info.panel_count, info.levels
        self.level_start = np.array([info.level_start[i] for i in range(self.levels + 1)], np.int32)

    def panels(self) -> np.ndarray:
        dt = PANEL3D_DTYPE if self.dim == 3 else PANEL2D_DTYPE
        out = np.zeros(self.panel_count, dt)
        f = lib().btk_tree_copy_panels_3d if self.dim == 3 else lib().btk_tree_copy_panels_2d
        _check(f(self.h, out.ctypes.data))
        return out

    def point_leaf_panel(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1), np.int32)
        _check(lib().btk_tree_copy_point_leaf_panel(self.h, out.ctypes.data))
        return out[: self.n]

    def panel_points_inside(self, width: int) -> np.ndarray:
        out = np.zeros((self.panel_count, width), np.int32)
        _check(lib().btk_tree_copy_panel_points_inside(self.h, width, out.ctypes.data))
        return out

    def permutation(self):
        perm = np.zeros(max(self.n, 1), np.int32)
        start = np.zeros(self.panel_count + 1, np.int32)
        _check(lib().btk_tree_copy_permutation(self.h, perm.ctypes.data, start.ctypes.data))
        return perm[: self.n], start

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and _lib is not None:
            _lib.btk_tree_destroy(h)
            self.h = None


class InteractionList:
    """A device-resident interaction list (btk_list handle).  learned_mac:
    the 3D traversal with the installed learned MAC, whose pairs carry their
    degree index in bits 2-3 of interact_type (degree 4 + 2 * index)."""

    def __init__(self, cfg: RunConfig, target: Tree, source: Tree, stream=None, learned_mac=False):
        h = C.c_void_p()
        if learned_mac:
            f = lib().btk_dual_tree_traversal_learned_3d
        else:
            f = lib().btk_dual_tree_traversal_3d if cfg.dim == 3 else lib().btk_dual_tree_traversal_2d
        _check(f(C.byref(cfg), target.h, source.h, _stream(stream), C.byref(h)))
        self.h = h
        self.count = lib().btk_list_count(self.h)

    def pairs(self) -> np.ndarray:
        out = np.zeros(self.count, PAIR_DTYPE)
        _check(lib().btk_list_copy(self.h, out.ctypes.data))
        return out

    def split(self, cfg: RunConfig):
        outs = [np.zeros(k, PAIR_DTYPE) for k in
                (cfg.fmm_pp_count, cfg.fmm_pc_count, cfg.fmm_cp_count, cfg.fmm_cc_count)]
        _check(lib().btk_split_interactions(self.h, *[o.ctypes.data for o in outs]))
        return tuple(outs)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and _lib is not None:
            _lib.btk_list_destroy(h)
            self.h = None


# ---- reference-named entry points ---------------------------------------

def blfmm_tree_construction_2d(cfg, x, y, stream=None) -> Tree:
    return Tree(cfg, x, y, None, stream)


def blfmm_tree_construction_3d(cfg, x, y, z, stream=None) -> Tree:
    return Tree(cfg, x, y, z, stream)


def dual_tree_traversal_2d(cfg, target: Tree, source: Tree, stream=None) -> InteractionList:
    return InteractionList(cfg, target, source, stream)


dual_tree_traversal_3d = dual_tree_traversal_2d


def split_interactions(lst: InteractionList, cfg: RunConfig):
    return lst.split(cfg)


def _upward(cfg, tree: Tree, charges, W, stream=None):
    f = lib().btk_upward_pass_3d if cfg.dim == 3 else lib().btk_upward_pass_2d
    _check(f(C.byref(cfg), tree.h, _ptr(charges), _ptr(W), _stream(stream)))
    return W


def upward_pass_2d(cfg, tree, charges, W, stream=None):
    return _upward(cfg, tree, charges, W, stream)


upward_pass_3d = upward_pass_2d

NOALLOC_MAX_POINTS = 8388608  # BTK_NOALLOC_MAX_POINTS


def upward_pass_3d_noalloc(cfg, tree, charges, W, stream=None):
    """btk_upward_pass_3d_noalloc: the upward pass with no device allocation
    (its scratch is a fixed array per GPU); the work stays queued on the
    stream."""
    _check(lib().btk_upward_pass_3d_noalloc(C.byref(cfg), tree.h, _ptr(charges), _ptr(W), _stream(stream)))
    return W


def _interactions(cfg, opts, target, source, lst, charges, W, soln, Psi, stream=None):
    f = lib().btk_fmm_interactions_3d if cfg.dim == 3 else lib().btk_fmm_interactions_2d
    _check(f(C.byref(cfg), C.byref(opts), target.h, source.h, lst.h, _ptr(charges), _ptr(W), _ptr(soln),
             _ptr(Psi), _stream(stream)))


def poisson_fmm_interactions_2d(cfg, target, source, lst, charges, W, soln, Psi, opts=None, stream=None):
    _interactions(cfg, opts or options(KERNEL_POT), target, source, lst, charges, W, soln, Psi, stream)


poisson_fmm_interactions_3d = poisson_fmm_interactions_2d


def poisson_vel_fmm_interactions_3d(cfg, target, source, lst, charges, W, vel, Psi3, opts=None, stream=None):
    o = opts or options(KERNEL_REF_VEL3D)
    o.kernel = KERNEL_REF_VEL3D
    _interactions(cfg, o, target, source, lst, charges, W, vel, Psi3, stream)


def _downward(cfg, opts, target, Psi, soln, stream=None):
    f = lib().btk_downward_pass_3d if cfg.dim == 3 else lib().btk_downward_pass_2d
    _check(f(C.byref(cfg), C.byref(opts), target.h, _ptr(Psi), _ptr(soln), _stream(stream)))


def downward_pass_2d(cfg, target, Psi, soln, opts=None, stream=None):
    _downward(cfg, opts or options(KERNEL_POT), target, Psi, soln, stream)


downward_pass_3d = downward_pass_2d


def downward_pass_3d_3(cfg, target, Psi3, vel, opts=None, stream=None):
    o = opts or options(KERNEL_REF_VEL3D)
    o.kernel = KERNEL_REF_VEL3D
    _downward(cfg, o, target, Psi3, vel, stream)


def direct_sum(dim, opts, eps, xt, yt, zt, xs, ys, zs, q, out, idx=None, stream=None):
    m = int(idx.numel()) if idx is not None else int(xt.numel())
    if dim == 3:
        _check(lib().btk_direct_sum_3d(C.byref(opts), float(eps), m, _ptr(idx), _ptr(xt), _ptr(yt), _ptr(zt),
                                       int(xs.numel()), _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q), _ptr(out),
                                       _stream(stream)))
    else:
        _check(lib().btk_direct_sum_2d(C.byref(opts), float(eps), m, _ptr(idx), _ptr(xt), _ptr(yt),
                                       int(xs.numel()), _ptr(xs), _ptr(ys), _ptr(q), _ptr(out), _stream(stream)))
    return out


def solve(cfg, opts, xt, yt, zt, q, out, xs=None, ys=None, zs=None, stream=None) -> SolveStats:
    """Whole pipeline (tree .. downward); adds into out[C][M]."""
    coincident = xs is None
    if coincident:
        xs, ys, zs = xt, yt, zt
    st = SolveStats()
    _check(lib().btk_solve(C.byref(cfg), C.byref(opts), int(xt.numel()), _ptr(xt), _ptr(yt), _ptr(zt),
                           int(xs.numel()), _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q), int(coincident), _ptr(out),
                           C.byref(st), _stream(stream)))
    return st


def autotune_multi(cfg, opts, xt, yt, zt, q, xs=None, ys=None, zs=None, stream=None) -> Tuning:
    """Times the many-vector MFMA launches of this problem (q is [K][n]) with
    32, 64 and 128 charge vectors per workgroup and installs the fastest
    process-wide (btk_autotune_multi)."""
    coincident = xs is None
    if coincident:
        xs, ys, zs = xt, yt, zt
    t = Tuning()
    _check(lib().btk_autotune_multi(C.byref(cfg), C.byref(opts), int(xt.numel()), _ptr(xt), _ptr(yt), _ptr(zt),
                                    int(xs.numel()), _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q), int(q.shape[0]),
                                    int(coincident), C.byref(t), _stream(stream)))
    return t


def solve_multi(cfg, opts, xt, yt, zt, q, out, xs=None, ys=None, zs=None, stream=None) -> SolveStats:
    """Many charge vectors over one geometry: q is [K][n], out [K][C][M] (added to)."""
    coincident = xs is None
    if coincident:
        xs, ys, zs = xt, yt, zt
    K = int(q.shape[0])
    st = SolveStats()
    _check(lib().btk_solve_multi(C.byref(cfg), C.byref(opts), int(xt.numel()), _ptr(xt), _ptr(yt), _ptr(zt),
                                 int(xs.numel()), _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q), K, int(coincident),
                                 _ptr(out), C.byref(st), _stream(stream)))
    return st


def solve_batch(cfg, opts, offsets, x, y, z, q, out, stream=None) -> SolveStats:
    """Many independent problems in one solve: problem p owns points
    offsets[p]:offsets[p+1]; adds into out[C][N]."""
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    st = SolveStats()
    _check(lib().btk_solve_batch(C.byref(cfg), C.byref(opts), len(off) - 1,
                                 off.ctypes.data_as(C.POINTER(C.c_int)), _ptr(x), _ptr(y), _ptr(z), _ptr(q),
                                 _ptr(out), C.byref(st), _stream(stream)))
    return st


class Plan:
    """A solve plan for one geometry (btk_plan_create): execute(q, out) runs
    the upward pass, interactions and downward pass for a new charge vector
    as one HIP graph launch, adding into out[C][M]."""

    def __init__(self, cfg, opts, xt, yt, zt, xs=None, ys=None, zs=None, stream=None):
        coincident = xs is None
        if coincident:
            xs, ys, zs = xt, yt, zt
        self._keep = (xt, yt, zt, xs, ys, zs)
        self.h = C.c_void_p()
        _check(lib().btk_plan_create(C.byref(cfg), C.byref(opts), int(xt.numel()), _ptr(xt), _ptr(yt), _ptr(zt),
                                     int(xs.numel()), _ptr(xs), _ptr(ys), _ptr(zs), int(coincident), _stream(stream),
                                     C.byref(self.h)))

    @classmethod
    def batch(cls, cfg, opts, offsets, x, y, z=None, stream=None):
        """A plan for a fixed forest (btk_plan_create_batch); problem p owns
        points offsets[p] .. offsets[p+1]."""
        self = cls.__new__(cls)
        off = np.ascontiguousarray(np.asarray(offsets), np.int32)
        self._keep = (x, y, z, off)
        self.h = C.c_void_p()
        _check(lib().btk_plan_create_batch(C.byref(cfg), C.byref(opts), len(off) - 1,
                                           off.ctypes.data_as(C.POINTER(C.c_int)), _ptr(x), _ptr(y), _ptr(z),
                                           _stream(stream), C.byref(self.h)))
        return self

    def execute(self, q, out, stream=None):
        _check(lib().btk_plan_execute(self.h, _ptr(q), _ptr(out), _stream(stream)))
        return out

    def close(self):
        if self.h:
            lib().btk_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Learned MAC (btk_mac_*)
# ---------------------------------------------------------------------------
MAC_FEATURES, MAC_OUTPUTS, MAC_WEIGHTS = 10, 12, 5644


def mac_sample(cfg, x, y, z, n_samples, seed, X, Y, pairs=None, stream=None) -> int:
    """Training samples from the treecode's own traversal of (x, y, z):
    X[n][10] features, Y[n][12] measured log10(errors)/4, optionally the
    (target, source) panels pairs[n][2]; returns the count."""
    cnt = C.c_int()
    _check(lib().btk_mac_sample(C.byref(cfg), int(x.numel()), _ptr(x), _ptr(y), _ptr(z), int(n_samples),
                                int(seed), _ptr(X), _ptr(Y), _ptr(pairs), C.byref(cnt), _stream(stream)))
    return cnt.value


def mac_train(X, Y, n, epochs=50, batch=1024, lr=2e-3, seed=1, reinit=False, stream=None) -> float:
    loss = C.c_float()
    _check(lib().btk_mac_train(_ptr(X), _ptr(Y), int(n), int(epochs), int(batch), float(lr), int(seed),
                               int(bool(reinit)), C.byref(loss), _stream(stream)))
    return loss.value


def mac_set_weights(w: np.ndarray) -> None:
    w = np.ascontiguousarray(w, dtype=np.float32)
    _check(lib().btk_mac_set_weights(w.ctypes.data_as(C.POINTER(C.c_float)), int(w.size)))


def mac_get_weights() -> np.ndarray:
    w = np.zeros(MAC_WEIGHTS, np.float32)
    _check(lib().btk_mac_get_weights(w.ctypes.data_as(C.POINTER(C.c_float)), MAC_WEIGHTS))
    return w


def mac_set_budget(rel_err: float) -> None:
    _check(lib().btk_mac_set_budget(float(rel_err)))


def mac_set_costs(pp=1.0, pc=1.0, cp=1.0, cc=1.0) -> None:
    """Relative cost of one PP, PC, CP and CC kernel evaluation for the
    learned MAC's choice (btk_mac_set_costs; all 1 = fewest evaluations)."""
    c = (C.c_float * 4)(pp, pc, cp, cc)
    _check(lib().btk_mac_set_costs(c))


def mac_forward(X, n, out, stream=None):
    _check(lib().btk_mac_forward(_ptr(X), int(n), _ptr(out), _stream(stream)))
    return out


def mac_gradients(X, Y, n, grad, stream=None) -> float:
    loss = C.c_float()
    _check(lib().btk_mac_gradients(_ptr(X), _ptr(Y), int(n), _ptr(grad), C.byref(loss), _stream(stream)))
    return loss.value
