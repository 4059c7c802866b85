"""ctypes wrapper of the CPU restatement in oracle/btk_oracle.c.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py import this module.  It is the checker the GPU path is
compared against, never part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

KERNEL_POT, KERNEL_REF_VEL3D, KERNEL_POT_GRAD = 0, 1, 2


class RunConfig(C.Structure):
    """Mirror of RunConfig (structs.hpp:6-21)."""
    _fields_ = [
        ("dim", C.c_int), ("interp_degree", C.c_int), ("interp_point_count", C.c_int),
        ("fmm_theta", C.c_double), ("fmm_cluster_thresh", C.c_int),
        ("fmm_interaction_count", C.c_int), ("fmm_pp_count", C.c_int), ("fmm_pc_count", C.c_int),
        ("fmm_cp_count", C.c_int), ("fmm_cc_count", C.c_int), ("ker_eps", C.c_double),
        ("mpi_p", C.c_int), ("mpi_id", C.c_int),
    ]


class Panel(C.Structure):
    _fields_ = [
        ("id", C.c_int), ("level", C.c_int), ("parent_id", C.c_int), ("child", C.c_int * 8),
        ("is_leaf", C.c_int), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("radius", C.c_double), ("point_count", C.c_int),
    ]


class Tree(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("panel_count", C.c_int), ("levels", C.c_int),
        ("level_start", C.POINTER(C.c_int)), ("panels", C.POINTER(Panel)),
        ("point_leaf_panel", C.POINTER(C.c_int)), ("leaf_start", C.POINTER(C.c_int)),
        ("leaf_points", C.POINTER(C.c_int)), ("n_points", C.c_int),
    ]


class Pair(C.Structure):
    _fields_ = [("target_panel", C.c_int), ("source_panel", C.c_int), ("interact_type", C.c_int)]


class Opts(C.Structure):
    _fields_ = [("kernel", C.c_int), ("convention", C.c_int), ("exclude_self", C.c_int),
                ("deepest_only", C.c_int), ("near_stride", C.c_int), ("far_stride", C.c_int),
                ("pass_stride", C.c_int)]


PAIR_DTYPE = np.dtype([("target_panel", "<i4"), ("source_panel", "<i4"), ("interact_type", "<i4")])
# Reference AoS layouts (structs.hpp:29-67), x86-64 alignment.
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


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


def use_library(path: str | None):
    """Switches every function of this module to the restatement built at
    `path` (None: the default build); returns the previous library.  The
    list tests use it to load builds with a planted mistake."""
    global _lib
    prev = _lib
    _lib = None if path is None else _load(path)
    return prev


def _load(path: str):
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    L.orc_read_run_config.argtypes = [C.c_char_p, C.POINTER(RunConfig)]
    L.orc_interp_vals_bli.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_int]
    L.orc_bli_points_shift.argtypes = [dp, C.c_double, C.c_double, C.c_int]
    L.orc_tree_build.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, C.POINTER(Tree)]
    L.orc_tree_free.argtypes = [C.POINTER(Tree)]
    L.orc_pack_panels_2d.argtypes = [C.POINTER(Tree), C.c_void_p]
    L.orc_pack_panels_3d.argtypes = [C.POINTER(Tree), C.c_void_p]
    L.orc_panel_points_inside.argtypes = [C.POINTER(Tree), C.c_int, ip]
    L.orc_dual_tree_traversal.argtypes = [C.POINTER(RunConfig), C.POINTER(Tree), C.POINTER(Tree),
                                          C.POINTER(C.POINTER(Pair))]
    L.orc_split_interactions.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 4
    L.orc_upward_pass.argtypes = [C.POINTER(RunConfig), C.POINTER(Tree), dp, dp, dp, dp, dp]
    L.orc_fmm_interactions.argtypes = [C.POINTER(RunConfig), C.POINTER(Opts), C.POINTER(Tree),
                                       C.POINTER(Tree), dp, dp, dp, dp, dp, dp, dp, dp, C.c_void_p,
                                       C.c_int, dp, dp]
    L.orc_downward_pass.argtypes = [C.POINTER(RunConfig), C.POINTER(Opts), C.POINTER(Tree), dp, dp,
                                    dp, dp, dp]
    L.orc_direct_sum.argtypes = [C.c_int, C.c_double, C.POINTER(Opts), C.c_int, ip, dp, dp, dp,
                                 C.c_int, dp, dp, dp, dp, dp]
    L.orc_solve.argtypes = [C.POINTER(RunConfig), C.POINTER(Opts), C.c_int, dp, dp, dp, C.c_int, dp,
                            dp, dp, dp, C.c_int, dp, dp, ip]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_threads.restype = C.c_int
    L.orc_components.argtypes = [C.c_int, C.c_int]
    L.orc_last_near_seconds.restype = C.c_double
    L.orc_last_near_fraction.restype = C.c_double
    L.orc_last_far_seconds.restype = C.c_double
    L.orc_last_far_fraction.restype = C.c_double
    L.orc_last_passes.argtypes = [dp, dp]
    L.free = C.CDLL(None).free
    L.free.argtypes = [C.c_void_p]
    return L


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def make_config(dim, degree, thresh, theta, eps) -> RunConfig:
    c = RunConfig()
    c.dim = dim
    c.interp_degree = degree
    c.interp_point_count = (degree + 1) ** dim
    c.fmm_theta = theta
    c.fmm_cluster_thresh = thresh
    c.ker_eps = eps
    c.mpi_p, c.mpi_id = 1, 0
    return c


def read_run_config(path: str, dim: int):
    c = RunConfig()
    c.dim = dim
    c.fmm_theta = 0.7   # structs.hpp:10
    c.ker_eps = -1.0    # structs.hpp:17
    rc = lib().orc_read_run_config(path.encode(), C.byref(c))
    return rc, c


def interp_vals_bli(x, lo, hi, n):
    out = np.zeros(n + 1)
    lib().orc_interp_vals_bli(_dp(out), float(x), float(lo), float(hi), int(n))
    return out


def bli_points_shift(lo, hi, n):
    out = np.zeros(n + 1)
    lib().orc_bli_points_shift(_dp(out), float(lo), float(hi), int(n))
    return out


class OracleTree:
    """Owns an orc_tree; exposes numpy views of its arrays."""

    def __init__(self, dim, thresh, x, y, z=None):
        self.dim = dim
        self.thresh = thresh
        self.x, self.y = np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64)
        self.z = np.ascontiguousarray(z, np.float64) if z is not None else None
        self.t = Tree()
        rc = lib().orc_tree_build(dim, thresh, len(self.x), _dp(self.x), _dp(self.y), _dp(self.z),
                                  C.byref(self.t))
        if rc != 0:
            raise RuntimeError(f"orc_tree_build failed: {rc}")
        pc, n = self.t.panel_count, self.t.n_points
        self.panel_count = pc
        self.levels = self.t.levels
        self.level_start = np.ctypeslib.as_array(self.t.level_start, (self.levels + 1,)).copy()
        self.point_leaf_panel = np.ctypeslib.as_array(self.t.point_leaf_panel, (max(n, 1),))[:n].copy()
        self.leaf_start = np.ctypeslib.as_array(self.t.leaf_start, (pc + 1,)).copy()
        self.leaf_points = np.ctypeslib.as_array(self.t.leaf_points, (max(n, 1),))[:n].copy()

    def packed_panels(self) -> np.ndarray:
        dt = PANEL3D_DTYPE if self.dim == 3 else PANEL2D_DTYPE
        out = np.zeros(self.panel_count, dtype=dt)
        f = lib().orc_pack_panels_3d if self.dim == 3 else lib().orc_pack_panels_2d
        f(C.byref(self.t), out.ctypes.data)
        return out

    def panel_points_inside(self) -> np.ndarray:
        out = np.zeros((self.panel_count, self.thresh), np.int32)
        lib().orc_panel_points_inside(C.byref(self.t), self.thresh, _ip(out))
        return out

    def __del__(self):
        try:
            lib().orc_tree_free(C.byref(self.t))
        except Exception:
            pass


def traversal(cfg: RunConfig, tt: OracleTree, ts: OracleTree):
    p = C.POINTER(Pair)()
    cnt = lib().orc_dual_tree_traversal(C.byref(cfg), C.byref(tt.t), C.byref(ts.t), C.byref(p))
    arr = np.zeros(cnt, dtype=PAIR_DTYPE)
    if cnt:
        C.memmove(arr.ctypes.data, p, cnt * 12)
    lib().free(C.cast(p, C.c_void_p))
    return arr


def split_interactions(lst: np.ndarray, cfg: RunConfig):
    pp = np.zeros(cfg.fmm_pp_count, PAIR_DTYPE)
    pc = np.zeros(cfg.fmm_pc_count, PAIR_DTYPE)
    cp = np.zeros(cfg.fmm_cp_count, PAIR_DTYPE)
    cc = np.zeros(cfg.fmm_cc_count, PAIR_DTYPE)
    lib().orc_split_interactions(lst.ctypes.data, len(lst), pp.ctypes.data, pc.ctypes.data,
                                 cp.ctypes.data, cc.ctypes.data)
    return pp, pc, cp, cc


def opts(kernel=KERNEL_POT, convention=0, exclude_self=0, deepest_only=0, near_stride=1, far_stride=1,
         pass_stride=1) -> Opts:
    o = Opts()
    o.kernel, o.convention, o.exclude_self, o.deepest_only = kernel, convention, exclude_self, deepest_only
    o.near_stride, o.far_stride, o.pass_stride = near_stride, far_stride, pass_stride
    return o


def last_near():
    """(seconds in the PP/PC loops, fraction of PP/PC evaluations performed) of the last call."""
    return lib().orc_last_near_seconds(), lib().orc_last_near_fraction()


def last_passes():
    """(seconds, processed fraction) of the last solve's anterpolation,
    child-to-parent, parent-to-child and leaf-to-point passes."""
    sec, frac = np.zeros(4), np.zeros(4)
    lib().orc_last_passes(_dp(sec), _dp(frac))
    return sec, frac


def last_far():
    """(seconds in the CP/CC loops, fraction of CP/CC evaluations performed) of the last call."""
    return lib().orc_last_far_seconds(), lib().orc_last_far_fraction()


def components(dim, kernel):
    return lib().orc_components(dim, kernel)


def upward_pass(cfg, ts: OracleTree, q):
    P = (cfg.interp_degree + 1) ** cfg.dim
    W = np.zeros(ts.panel_count * P)
    q = np.ascontiguousarray(q, np.float64)
    lib().orc_upward_pass(C.byref(cfg), C.byref(ts.t), _dp(ts.x), _dp(ts.y), _dp(ts.z), _dp(q), _dp(W))
    return W.reshape(ts.panel_count, P)


def fmm_interactions(cfg, o: Opts, tt: OracleTree, ts: OracleTree, q, W, lst):
    P = (cfg.interp_degree + 1) ** cfg.dim
    Cn = components(cfg.dim, o.kernel)
    M = len(tt.x)
    out = np.zeros(Cn * M)
    Psi = np.zeros(Cn * tt.panel_count * P)
    q = np.ascontiguousarray(q, np.float64)
    W = np.ascontiguousarray(W, np.float64).ravel()
    lst = np.ascontiguousarray(lst)
    lib().orc_fmm_interactions(C.byref(cfg), C.byref(o), C.byref(tt.t), C.byref(ts.t), _dp(tt.x), _dp(tt.y),
                               _dp(tt.z), _dp(ts.x), _dp(ts.y), _dp(ts.z), _dp(q), _dp(W), lst.ctypes.data,
                               len(lst), _dp(out), _dp(Psi))
    return out.reshape(Cn, M), Psi.reshape(Cn, tt.panel_count, P)


def downward_pass(cfg, o: Opts, tt: OracleTree, Psi, out):
    Psi = np.ascontiguousarray(Psi, np.float64).copy()
    out = np.ascontiguousarray(out, np.float64).copy()
    lib().orc_downward_pass(C.byref(cfg), C.byref(o), C.byref(tt.t), _dp(tt.x), _dp(tt.y), _dp(tt.z),
                            _dp(Psi.ravel()), _dp(out.ravel()))
    return out


def direct_sum(dim, eps, o: Opts, xt, yt, zt, xs, ys, zs, q, idx=None):
    Cn = components(dim, o.kernel)
    m = len(xt) if idx is None else len(idx)
    out = np.zeros(Cn * m)
    idx_a = None if idx is None else np.ascontiguousarray(idx, np.int32)
    f = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
    xt, yt, zt, xs, ys, zs, q = map(f, (xt, yt, zt, xs, ys, zs, q))
    lib().orc_direct_sum(dim, float(eps), C.byref(o), m, _ip(idx_a), _dp(xt), _dp(yt), _dp(zt), len(xs),
                         _dp(xs), _dp(ys), _dp(zs), _dp(q), _dp(out))
    return out.reshape(Cn, m)


def solve(cfg, o: Opts, xt, yt, zt, xs, ys, zs, q, coincident=True):
    """Whole CPU pipeline; returns (out[C][M], phase seconds[5], counts[4])."""
    Cn = components(cfg.dim, o.kernel)
    f = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
    xt, yt, zt, xs, ys, zs, q = map(f, (xt, yt, zt, xs, ys, zs, q))
    M = len(xt)
    out = np.zeros(Cn * M)
    t = np.zeros(5)
    counts = np.zeros(4, np.int32)
    rc = lib().orc_solve(C.byref(cfg), C.byref(o), M, _dp(xt), _dp(yt), _dp(zt), len(xs), _dp(xs), _dp(ys),
                         _dp(zs), _dp(q), int(coincident), _dp(out), _dp(t), _ip(counts))
    if rc != 0:
        raise RuntimeError(f"orc_solve failed: {rc}")
    return out.reshape(Cn, M), t, counts
