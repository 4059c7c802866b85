"""Pins the oracle's tree construction and dual traversal (tree_construction.cpp,
interaction_list.cpp) without the reference binary:

- hand-derived tiny trees (tests/golden/tiny_trees.json);
- invariants of every tree: boxes are exact halves, every point lies in its
  leaf's box on the side the strict '>' rule puts it, leaf lists partition the
  points in ascending order, counts add up, ids are level-sorted, a panel is
  refined iff it holds more than cluster_thresh points;
- the root-box quirk (numeric_limits<double>::min(), tree_construction.cpp:146);
- brute-force coverage of the traversal: every (target point, source point)
  pair is covered by exactly one list entry, every far-field entry satisfies
  the MAC, and the types follow the point counts (interaction_list.cpp:155-177).
"""
import json
import os
import sys

import numpy as np
import pytest

from barytreek_amd import workloads as wl
from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "tiny_trees.json")
DBL_MIN = sys.float_info.min


def panel_arrays(t: O.OracleTree):
    p = t.t.panels
    pc = t.panel_count
    lo = np.array([[p[i].lo[d] for d in range(3)] for i in range(pc)])
    hi = np.array([[p[i].hi[d] for d in range(3)] for i in range(pc)])
    child = np.array([[p[i].child[c] for c in range(8)] for i in range(pc)])
    par = np.array([p[i].parent_id for i in range(pc)])
    cnt = np.array([p[i].point_count for i in range(pc)])
    lvl = np.array([p[i].level for i in range(pc)])
    leaf = np.array([p[i].is_leaf for i in range(pc)], bool)
    rad = np.array([p[i].radius for i in range(pc)])
    return lo, hi, child, par, cnt, lvl, leaf, rad


@pytest.mark.parametrize("case", json.load(open(GOLD))["trees"], ids=lambda c: c["name"])
def test_hand_derived_trees(case):
    dim = case["dim"]
    x, y = np.array(case["x"]), np.array(case["y"])
    z = np.array(case["z"]) if dim == 3 else None
    t = O.OracleTree(dim, case["thresh"], x, y, z)
    lo, hi, child, par, cnt, lvl, leaf, rad = panel_arrays(t)
    assert t.panel_count == case["panel_count"]
    assert list(t.level_start) == case["level_start"]
    assert list(t.point_leaf_panel) == case["point_leaf_panel"]
    assert list(par) == case["parent"]
    assert list(child[:, 0]) == case["first_child"]
    assert list(cnt) == case["point_count"]
    for k, pts in case["leaf_points"].items():
        k = int(k)
        assert list(t.leaf_points[t.leaf_start[k]:t.leaf_start[k + 1]]) == pts
    for k, box in case["boxes"].items():
        k = int(k)
        for d in range(dim):
            assert (lo[k, d], hi[k, d]) == tuple(box[d])


TREE_INPUTS = [
    ("cube3", 3, 5000, 60),
    ("sphere3", 3, 6000, 50),
    ("gauss2", 2, 8000, 40),
    ("cube2", 2, 3000, 7),
]


def make_points(kind, n, seed):
    if kind == "cube3":
        return wl.uniform_cube(n, 3, seed)
    if kind == "cube2":
        return wl.uniform_cube(n, 2, seed)
    if kind == "gauss2":
        return wl.gaussian_mixture_2d(n, seed)
    return wl.sphere_plummer_3d(n, seed)


@pytest.mark.parametrize("kind,dim,n,thresh", TREE_INPUTS)
def test_tree_invariants(kind, dim, n, thresh):
    x, y, z, _ = make_points(kind, n, 5)
    t = O.OracleTree(dim, thresh, x, y, z)
    lo, hi, child, par, cnt, lvl, leaf, rad = panel_arrays(t)
    nc = 2 ** dim
    co = [x, y, z]
    # root box: data min/max
    for d in range(dim):
        assert lo[0, d] == co[d].min() and hi[0, d] == co[d].max()
    # level-sorted ids and level_start
    assert np.all(np.diff(lvl) >= 0)
    for l in range(t.levels):
        assert np.all(lvl[t.level_start[l]:t.level_start[l + 1]] == l)
    # refine iff count > thresh; counts add up; children are exact halves in c1..c8 order
    hx, hy, hz = [1, 0, 0, 1, 1, 0, 0, 1], [1, 1, 0, 0, 1, 1, 0, 0], [1, 1, 1, 1, 0, 0, 0, 0]
    for p in range(t.panel_count):
        assert leaf[p] == (cnt[p] <= thresh)
        if leaf[p]:
            continue
        kids = child[p, :nc]
        assert np.all(np.diff(kids) == 1) and np.all(par[kids] == p)
        assert cnt[kids].sum() == cnt[p]
        mid = 0.5 * (lo[p] + hi[p])
        for c, k in enumerate(kids):
            for d, h in zip(range(dim), (hx[c], hy[c], hz[c])):
                assert (lo[k, d], hi[k, d]) == ((mid[d], hi[p, d]) if h else (lo[p, d], mid[d]))
    # radius: half the box diagonal
    ext = hi[:, :dim] - lo[:, :dim]
    assert np.allclose(rad, 0.5 * np.sqrt((ext ** 2).sum(1)), rtol=1e-15, atol=0)
    # leaf lists: ascending, a partition of 0..n-1, consistent with point_leaf_panel
    allpts = []
    for p in np.nonzero(leaf)[0]:
        pts = t.leaf_points[t.leaf_start[p]:t.leaf_start[p + 1]]
        assert len(pts) == cnt[p] and np.all(np.diff(pts) > 0)
        assert np.all(t.point_leaf_panel[pts] == p)
        allpts.append(pts)
        # every point sits on the side of each ancestor split that '>' gives it
        for d in range(dim):
            v = co[d][pts]
            assert np.all(v >= lo[p, d]) and np.all(v <= hi[p, d])
            if p != 0:
                q = par[p]
                mid = 0.5 * (lo[q, d] + hi[q, d])
                upper = lo[p, d] == mid
                assert np.all(v > mid) if upper else np.all(v <= mid)
    assert np.array_equal(np.sort(np.concatenate(allpts)), np.arange(n))


def test_root_max_uses_dbl_min_on_negative_axis():
    rng = np.random.default_rng(3)
    x = -rng.uniform(0.5, 1.0, 500)
    y = rng.uniform(-1, 1, 500)
    t = O.OracleTree(2, 50, x, y)
    lo, hi, *_ = panel_arrays(t)
    assert hi[0, 0] == DBL_MIN and lo[0, 0] == x.min()


def test_root_is_leaf_when_small():
    x, y, z, _ = wl.uniform_cube(10, 3, 1)
    t = O.OracleTree(3, 10, x, y, z)
    assert t.panel_count == 1 and t.levels == 1 and list(t.level_start) == [0, 1]


def subtree_points(t: O.OracleTree, p, child, leaf, nc):
    if leaf[p]:
        return t.leaf_points[t.leaf_start[p]:t.leaf_start[p + 1]]
    return np.concatenate([subtree_points(t, c, child, leaf, nc) for c in child[p, :nc]])


@pytest.mark.parametrize("dim,n,thresh,theta,split", [(3, 1500, 40, 0.7, False), (2, 1800, 30, 0.6, False),
                                                      (2, 1200, 25, 0.5, True), (3, 1000, 60, 0.9, True)])
def test_traversal_covers_every_pair_once(dim, n, thresh, theta, split):
    xt, yt, zt, _ = wl.uniform_cube(n, dim, 21)
    if split:
        xs, ys, zs, _ = (wl.gaussian_mixture_2d(n + 100, 22) if dim == 2 else wl.sphere_plummer_3d(n + 100, 22))
    else:
        xs, ys, zs = xt, yt, zt
    tt = O.OracleTree(dim, thresh, xt, yt, zt)
    ts = O.OracleTree(dim, thresh, xs, ys, zs) if split else tt
    cfg = O.make_config(dim, 4, thresh, theta, 0.0)
    lst = O.traversal(cfg, tt, ts)
    nc = 2 ** dim
    lo_t, hi_t, child_t, _, cnt_t, _, leaf_t, rad_t = panel_arrays(tt)
    lo_s, hi_s, child_s, _, cnt_s, _, leaf_s, rad_s = panel_arrays(ts)
    cover = np.zeros((len(xt), len(xs)), np.int32)
    counts = [0, 0, 0, 0]
    for e in lst:
        T, S, ty = int(e["target_panel"]), int(e["source_panel"]), int(e["interact_type"])
        counts[ty] += 1
        # types follow the point counts (F6)
        assert ty == 2 * (cnt_t[T] > thresh) + (cnt_s[S] > thresh)
        ct = 0.5 * (lo_t[T, :dim] + hi_t[T, :dim])
        cs = 0.5 * (lo_s[S, :dim] + hi_s[S, :dim])
        dist = np.linalg.norm(cs - ct)
        mac = dist > 0 and (rad_t[T] + rad_s[S]) / dist < theta
        if ty != 0:
            assert mac
        else:
            assert mac or (leaf_t[T] and leaf_s[S])
        ti = subtree_points(tt, T, child_t, leaf_t, nc)
        si = subtree_points(ts, S, child_s, leaf_s, nc)
        cover[np.ix_(ti, si)] += 1
    assert np.all(cover == 1)
    assert counts == [cfg.fmm_pp_count, cfg.fmm_pc_count, cfg.fmm_cp_count, cfg.fmm_cc_count]
    assert cfg.fmm_interaction_count == len(lst)
    pp, pc, cp, cc = O.split_interactions(lst, cfg)
    for ty, part in enumerate((pp, pc, cp, cc)):
        assert np.array_equal(part, lst[lst["interact_type"] == ty])
