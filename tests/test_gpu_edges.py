"""Edge cases of the path against the oracle: empty and tiny inputs, points
on the splitting planes (the strict '>' rule), more coincident points than a
leaf holds, the lowest and highest interpolation degrees, and a MAC that
accepts nothing (the treecode is then the direct sum); the solve with separate
target and source points and with the reference's leaf set for leaf-to-point."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 1e-12


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def tree_and_list_match(dim, thresh, x, y, z):
    """GPU tree and list against the oracle's, byte for byte."""
    cfg = btk.make_config(dim, 4, thresh, 0.7, 1e-4)
    t = btk.Tree(cfg, dev(x), dev(y), dev(z))
    ot = O.OracleTree(dim, thresh, x, y, z)
    assert t.panel_count == ot.panel_count and t.levels == ot.levels
    assert np.array_equal(t.level_start, ot.level_start)
    assert t.panels().tobytes() == ot.packed_panels().tobytes(), "panels differ"
    if len(x):
        assert np.array_equal(t.point_leaf_panel(), ot.point_leaf_panel)
    ocfg = O.make_config(dim, 4, thresh, 0.7, 1e-4)
    olist = O.traversal(ocfg, ot, ot)
    lst = btk.InteractionList(cfg, t, t)
    assert lst.pairs().tobytes() == olist.tobytes(), "interaction lists differ"
    return t.panel_count, len(olist)


def solve_matches_oracle(dim, deg, thresh, theta, eps, kernel, excl, x, y, z, q):
    n = len(x)
    cfg = btk.make_config(dim, deg, thresh, theta, eps)
    Cn = btk.components(dim, kernel)
    out = torch.zeros(max(Cn * n, 1), dtype=torch.float64, device="cuda")
    st = btk.solve(cfg, btk.options(kernel=kernel, exclude_self=excl), dev(x), dev(y), dev(z), dev(q), out)
    O.set_threads(16)
    oout, _, ocnt = O.solve(O.make_config(dim, deg, thresh, theta, eps), O.opts(kernel=kernel, exclude_self=excl),
                            x, y, z, x, y, z, q)
    assert list(st.count) == list(ocnt)
    g = out.cpu().numpy()[:Cn * n]
    e = rel(g, oout)
    return e, g


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("n", [0, 1, 2, 7])
def test_empty_and_tiny_inputs(dim, n):
    print(f"\n[edges] dim={dim} n={n}", flush=True)
    x, y, z, q = wl.uniform_cube(n, dim, 40 + n)
    pcount, npairs = tree_and_list_match(dim, 4, x, y, z)
    print(f"  panels {pcount}, pairs {npairs}", flush=True)
    for eps, excl in ((1e-3, 0), (0.0, 1)):
        e, g = solve_matches_oracle(dim, 4, 4, 0.7, eps, btk.KERNEL_POT_GRAD, excl, x, y, z, q)
        print(f"  eps={eps} exclude_self={excl}: rel {e:.2e}", flush=True)
        assert e < TOL
        if n == 1 and excl:
            assert np.all(g == 0.0)  # the lone point's own term is excluded


@pytest.mark.parametrize("dim", [2, 3])
def test_points_on_splitting_planes(dim):
    """A lattice whose coordinates include every box midpoint down to level 4:
    each point on a plane goes to the lower half (strict '>',
    tree_construction.cpp:201-216)."""
    k = 17 if dim == 3 else 65
    g = np.linspace(-1.0, 1.0, k)
    mesh = np.meshgrid(*([g] * dim), indexing="ij")
    x, y = mesh[0].ravel(), mesh[1].ravel()
    z = mesh[2].ravel() if dim == 3 else None
    q = np.random.default_rng(3).uniform(-1, 1, len(x))
    print(f"\n[edges] lattice dim={dim} n={len(x)}", flush=True)
    pcount, npairs = tree_and_list_match(dim, 40, x, y, z)
    print(f"  panels {pcount}, pairs {npairs}", flush=True)
    e, _ = solve_matches_oracle(dim, 5, 40, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1, x, y, z, q)
    print(f"  solve rel {e:.2e}", flush=True)
    assert e < TOL


def test_coincident_points_within_a_leaf():
    """Groups of identical points no larger than a leaf: every zero-distance
    pair is excluded, as in the oracle."""
    rng = np.random.default_rng(8)
    base = rng.uniform(-1, 1, (400, 3))
    pts = np.repeat(base, 5, axis=0)  # 5 copies of each point
    x, y, z = pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy()
    q = rng.uniform(-1, 1, len(x))
    print(f"\n[edges] coincident groups n={len(x)}", flush=True)
    tree_and_list_match(3, 50, x, y, z)
    e, g = solve_matches_oracle(3, 6, 50, 0.7, 0.0, btk.KERNEL_POT, 1, x, y, z, q)
    print(f"  solve rel {e:.2e}", flush=True)
    assert e < TOL and np.all(np.isfinite(g))


def test_more_coincident_points_than_a_leaf_is_an_error():
    """50 identical points with leaf size 10 can never be split: the reference
    recurses without end; the oracle and the library stop at 200 levels with
    an error."""
    x = np.full(50, 0.25)
    print("\n[edges] 50 identical points, leaf 10", flush=True)
    with pytest.raises(RuntimeError):
        O.OracleTree(3, 10, x, x, x)
    cfg = btk.make_config(3, 4, 10, 0.7, 0.0)
    with pytest.raises(btk.BtkError, match="levels"):
        btk.Tree(cfg, dev(x), dev(x), dev(x))
    # the library stays usable afterwards
    xs, ys, zs, qs = wl.uniform_cube(2000, 3, 9)
    e, _ = solve_matches_oracle(3, 4, 100, 0.7, 0.0, btk.KERNEL_POT, 1, xs, ys, zs, qs)
    assert e < TOL


@pytest.mark.parametrize("dim,deg", [(3, 1), (2, 1), (2, 10)])
def test_lowest_and_highest_degrees(dim, deg):
    x, y, z, q = wl.uniform_cube(8000, dim, 50 + deg)
    print(f"\n[edges] dim={dim} degree={deg}", flush=True)
    e, _ = solve_matches_oracle(dim, deg, 150, 0.6, 1e-4, btk.KERNEL_POT_GRAD, 0, x, y, z, q)
    print(f"  solve rel {e:.2e}", flush=True)
    assert e < TOL


@pytest.mark.parametrize("dim", [2, 3])
def test_mac_accepting_nothing_is_the_direct_sum(dim):
    """theta = 0.01: no pair of boxes in the unit cube passes the MAC, so the
    list holds only PP pairs of leaves and the solve equals the direct sum."""
    n = 6000
    x, y, z, q = wl.uniform_cube(n, dim, 60 + dim)
    cfg = btk.make_config(dim, 4, 200, 0.01, 0.0)
    o = btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1)
    Cn = btk.components(dim, btk.KERNEL_POT_GRAD)
    out = torch.zeros(Cn * n, dtype=torch.float64, device="cuda")
    st = btk.solve(cfg, o, dev(x), dev(y), dev(z), dev(q), out)
    assert st.count[1] == st.count[2] == st.count[3] == 0 and st.count[0] > 0
    ds = torch.zeros_like(out)
    btk.direct_sum(dim, o, 0.0, dev(x), dev(y), dev(z), dev(x), dev(y), dev(z), dev(q), ds)
    e = float(torch.linalg.norm(out - ds) / torch.linalg.norm(ds))
    print(f"\n[edges] theta 0.01 dim={dim}: {int(st.count[0])} PP pairs, rel vs direct sum {e:.2e}", flush=True)
    assert e < 1e-13


@pytest.mark.parametrize("dim", [2, 3])
def test_solve_with_separate_targets_matches_oracle(dim):
    """btk_solve with distinct target and source points (two trees,
    example_poisson_2d_split.cpp:80-170): the results, accumulated in the
    target tree's sorted order and moved to the caller's, equal the oracle's."""
    xt, yt, zt, _ = wl.uniform_cube(9000, dim, 70 + dim)
    xs, ys, zs, q = wl.uniform_cube(7000, dim, 80 + dim)
    xs, ys = 0.8 * xs + 0.1, 0.8 * ys - 0.1
    cfg = btk.make_config(dim, 6, 120, 0.7, 1e-3)
    o = btk.options(kernel=btk.KERNEL_POT_GRAD)
    Cn = btk.components(dim, btk.KERNEL_POT_GRAD)
    out = torch.zeros(Cn * len(xt), dtype=torch.float64, device="cuda")
    btk.solve(cfg, o, dev(xt), dev(yt), dev(zt), dev(q), out, xs=dev(xs), ys=dev(ys), zs=dev(zs))
    O.set_threads(16)
    oout, _, _ = O.solve(O.make_config(dim, 6, 120, 0.7, 1e-3), O.opts(kernel=O.KERNEL_POT_GRAD), xt, yt, zt,
                         xs, ys, zs, q, coincident=False)
    e = rel(out.cpu().numpy(), oout)
    print(f"\n[edges] separate targets dim={dim}: rel {e:.2e}", flush=True)
    assert e < TOL


def test_solve_deepest_only_matches_oracle():
    """btk_solve with the reference's leaf set for leaf-to-point (F5,
    downward_pass.cpp:14-16) on a tree of uneven depth."""
    x, y, z, q = wl.sphere_plummer_3d(12000, 91)
    cfg = btk.make_config(3, 5, 150, 0.7, 1e-4)
    o = btk.options(kernel=btk.KERNEL_POT)
    o.deepest_only = 1
    out = torch.zeros(len(x), dtype=torch.float64, device="cuda")
    btk.solve(cfg, o, dev(x), dev(y), dev(z), dev(q), out)
    oo = O.opts(kernel=O.KERNEL_POT)
    oo.deepest_only = 1
    O.set_threads(16)
    oout, _, _ = O.solve(O.make_config(3, 5, 150, 0.7, 1e-4), oo, x, y, z, x, y, z, q)
    e = rel(out.cpu().numpy(), oout)
    print(f"\n[edges] deepest_only solve: rel {e:.2e}", flush=True)
    assert e < TOL
