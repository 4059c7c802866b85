"""GPU parity: the HIP path against the CPU restatement (oracle), stage by stage.

Trees, interaction lists and counters must be bit-exact; proxy weights,
potentials and gradients within 1e-12 relative (FP64 path).  Every stage is
fed the oracle's own inputs for that stage, so a mismatch names the stage.
"""
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


def points(kind, n, seed):
    if kind == "cube3":
        return wl.uniform_cube(n, 3, seed)
    if kind == "cube2":
        return wl.uniform_cube(n, 2, seed, charges="normal")
    if kind == "gauss2":
        return wl.gaussian_mixture_2d(n, seed)
    if kind == "sphere3":
        return wl.sphere_plummer_3d(n, seed)
    if kind == "negx3":  # every x negative: root max_x is DBL_MIN (tree_construction.cpp:146)
        x, y, z, q = wl.uniform_cube(n, 3, seed)
        return -np.abs(x) - 0.5, y, z, q
    raise ValueError(kind)


TREE_CASES = [
    ("cube3", 3, 20000, 500, 1),
    ("cube3", 3, 100000, 500, 2),     # the 100k/leaf-500 problem: uniform depth
    ("sphere3", 3, 30000, 200, 3),
    ("negx3", 3, 5000, 100, 4),
    ("cube3", 3, 300, 500, 5),        # root is a leaf
    ("gauss2", 2, 40000, 300, 6),
    ("cube2", 2, 20000, 100, 7),
]


@pytest.mark.parametrize("kind,dim,n,thresh,seed", TREE_CASES)
def test_tree_and_lists_bit_exact(kind, dim, n, thresh, seed):
    print(f"\n[tree] {kind} dim={dim} n={n} thresh={thresh} seed={seed}", flush=True)
    x, y, z, q = points(kind, n, seed)
    cfg = btk.make_config(dim, 6, thresh, 0.7, 1e-4)
    t = btk.Tree(cfg, dev(x), dev(y), dev(z))
    ot = O.OracleTree(dim, thresh, x, y, z)
    assert t.panel_count == ot.panel_count and t.levels == ot.levels
    assert np.array_equal(t.level_start, ot.level_start)
    gp, op = t.panels(), ot.packed_panels()
    assert gp.tobytes() == op.tobytes(), "panels differ"
    assert np.array_equal(t.point_leaf_panel(), ot.point_leaf_panel)
    assert np.array_equal(t.panel_points_inside(thresh), ot.panel_points_inside())
    # traversal (interaction_list.cpp): FIFO order and counters
    ocfg = O.make_config(dim, 6, thresh, 0.7, 1e-4)
    olist = O.traversal(ocfg, ot, ot)
    lst = btk.InteractionList(cfg, t, t)
    assert lst.count == len(olist)
    assert lst.pairs().tobytes() == olist.tobytes(), "interaction lists differ"
    for f in ("fmm_interaction_count", "fmm_pp_count", "fmm_pc_count", "fmm_cp_count", "fmm_cc_count"):
        assert getattr(cfg, f) == getattr(ocfg, f), f
    for a, b in zip(lst.split(cfg), O.split_interactions(olist, ocfg)):
        assert a.tobytes() == b.tobytes()


def test_split_trees_bit_exact():
    """Distinct target and source trees (example_poisson_2d_split.cpp:80-93)."""
    print("\n[tree] split 2D", flush=True)
    xt, yt, _, _ = wl.uniform_cube(15000, 2, 11)
    xs, ys, _, q = wl.gaussian_mixture_2d(12000, 12)
    cfg = btk.make_config(2, 5, 200, 0.6, 1e-4)
    tt, ts = btk.Tree(cfg, dev(xt), dev(yt)), btk.Tree(cfg, dev(xs), dev(ys))
    ott, ots = O.OracleTree(2, 200, xt, yt), O.OracleTree(2, 200, xs, ys)
    ocfg = O.make_config(2, 5, 200, 0.6, 1e-4)
    assert btk.InteractionList(cfg, tt, ts).pairs().tobytes() == O.traversal(ocfg, ott, ots).tobytes()


STAGE_CASES = [
    # kind, dim, n, degree, thresh, theta, eps, kernel, exclude_self
    ("cube3", 3, 20000, 8, 500, 0.7, 1e-4, btk.KERNEL_POT, 0),
    ("cube3", 3, 12000, 6, 200, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1),
    ("sphere3", 3, 12000, 5, 150, 0.7, 1e-4, btk.KERNEL_REF_VEL3D, 0),
    ("gauss2", 2, 30000, 6, 300, 0.6, 0.0, btk.KERNEL_POT_GRAD, 1),
    ("cube2", 2, 20000, 8, 200, 0.6, 1e-4, btk.KERNEL_POT, 0),
    ("cube3", 3, 6000, 10, 300, 0.7, 1e-4, btk.KERNEL_POT, 0),
    # sphere + Plummer potential+gradient at the degrees 8 and 10
    ("sphere3", 3, 24000, 8, 400, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1),
    ("sphere3", 3, 20000, 10, 600, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1),
]


@pytest.mark.parametrize("kind,dim,n,deg,thresh,theta,eps,kernel,excl", STAGE_CASES)
def test_passes_match_oracle(kind, dim, n, deg, thresh, theta, eps, kernel, excl):
    print(f"\n[passes] {kind} dim={dim} n={n} deg={deg} kernel={kernel} excl={excl}", flush=True)
    x, y, z, q = points(kind, n, 100 + n)
    O.set_threads(16)
    cfg = btk.make_config(dim, deg, thresh, theta, eps)
    ocfg = O.make_config(dim, deg, thresh, theta, eps)
    opts = btk.options(kernel=kernel, exclude_self=excl)
    oopts = O.opts(kernel=kernel, exclude_self=excl)
    P = (deg + 1) ** dim
    Cn = btk.components(dim, kernel)
    t = btk.Tree(cfg, dev(x), dev(y), dev(z))
    ot = O.OracleTree(dim, thresh, x, y, z)
    lst = btk.InteractionList(cfg, t, t)
    olist = O.traversal(ocfg, ot, ot)
    dq = dev(q)
    # upward pass
    W = torch.zeros(t.panel_count * P, dtype=torch.float64, device="cuda")
    btk.upward_pass_3d(cfg, t, dq, W)
    oW = O.upward_pass(ocfg, ot, q)
    torch.cuda.synchronize()
    e = rel(W.cpu().numpy(), oW)
    print(f"  upward rel {e:.2e}", flush=True)
    assert e < TOL
    # interactions, fed the oracle's weights
    soln = torch.zeros(Cn * n, dtype=torch.float64, device="cuda")
    Psi = torch.zeros(Cn * t.panel_count * P, dtype=torch.float64, device="cuda")
    btk._interactions(cfg, opts, t, t, lst, dq, dev(oW.ravel()), soln, Psi)
    oout, oPsi = O.fmm_interactions(ocfg, oopts, ot, ot, q, oW, olist)
    e1, e2 = rel(soln.cpu().numpy(), oout), rel(Psi.cpu().numpy(), oPsi)
    print(f"  interactions rel out {e1:.2e} psi {e2:.2e}", flush=True)
    assert e1 < TOL and e2 < TOL
    # downward pass, fed the oracle's target proxies and near-field result
    for deepest in (0, 1):
        opts.deepest_only = deepest
        oopts.deepest_only = deepest
        s2 = dev(oout.ravel().copy())
        btk._downward(cfg, opts, t, dev(oPsi.ravel().copy()), s2)
        ofin = O.downward_pass(ocfg, oopts, ot, oPsi, oout)
        e3 = rel(s2.cpu().numpy(), ofin)
        print(f"  downward(deepest_only={deepest}) rel {e3:.2e}", flush=True)
        assert e3 < TOL


@pytest.mark.parametrize("kind,dim,n,deg,thresh,theta,eps,kernel,excl", STAGE_CASES[:4])
def test_solve_matches_oracle_and_direct_sum(kind, dim, n, deg, thresh, theta, eps, kernel, excl):
    print(f"\n[solve] {kind} dim={dim} n={n} deg={deg} kernel={kernel}", flush=True)
    x, y, z, q = points(kind, n, 200 + n)
    O.set_threads(16)
    cfg = btk.make_config(dim, deg, thresh, theta, eps)
    ocfg = O.make_config(dim, deg, thresh, theta, eps)
    opts = btk.options(kernel=kernel, exclude_self=excl)
    oopts = O.opts(kernel=kernel, exclude_self=excl)
    Cn = btk.components(dim, kernel)
    out = torch.zeros(Cn * n, dtype=torch.float64, device="cuda")
    st = btk.solve(cfg, opts, dev(x), dev(y), dev(z), dev(q), out)
    oout, _, ocnt = O.solve(ocfg, oopts, x, y, z, x, y, z, q)
    g = out.cpu().numpy()
    e = rel(g, oout)
    assert list(st.count) == list(ocnt)
    # FP64 direct sum on the GPU, checked against the oracle on a sample
    ds = torch.zeros(Cn * n, dtype=torch.float64, device="cuda")
    btk.direct_sum(dim, opts, eps, dev(x), dev(y), dev(z), dev(x), dev(y), dev(z), dev(q), ds)
    dsn = ds.cpu().numpy().reshape(Cn, n)
    idx = np.arange(0, n, 97, dtype=np.int32)
    ods = O.direct_sum(dim, eps, oopts, x, y, z, x, y, z, q, idx=idx)
    e_ds = rel(dsn[:, idx], ods)
    l2 = rel(g, dsn)
    print(f"  solve vs oracle {e:.2e}; direct sum vs oracle {e_ds:.2e}; treecode rel-L2 vs direct sum {l2:.2e}",
          flush=True)
    assert e < TOL
    assert e_ds < TOL
    assert l2 < 1e-4


MIXED_CASES = [
    ("gauss2", 2, 60000, 6, 500, 0.6, 0.0, btk.KERNEL_POT_GRAD, 1),
    ("cube2", 2, 30000, 6, 300, 0.6, 1e-4, btk.KERNEL_POT, 0),
    ("cube3", 3, 40000, 8, 500, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1),
    ("sphere3", 3, 30000, 6, 300, 0.7, 1e-4, btk.KERNEL_POT, 0),
    ("cube3", 3, 20000, 6, 300, 0.7, 1e-4, btk.KERNEL_REF_VEL3D, 0),
]

MIXED_TOL = 1e-6  # FP32 evaluation, FP64 accumulation: rel-L2 distance to the FP64 path


@pytest.mark.parametrize("kind,dim,n,deg,thresh,theta,eps,kernel,excl", MIXED_CASES)
def test_mixed_precision_close_to_fp64(kind, dim, n, deg, thresh, theta, eps, kernel, excl):
    print(f"\n[mixed] {kind} dim={dim} n={n} deg={deg} kernel={kernel}", flush=True)
    x, y, z, q = points(kind, n, 300 + n)
    cfg = btk.make_config(dim, deg, thresh, theta, eps)
    Cn = btk.components(dim, kernel)
    d = [dev(a) for a in (x, y, z, q)]
    res = {}
    for prec in (btk.PREC_FP64, btk.PREC_MIXED):
        opts = btk.options(kernel=kernel, exclude_self=excl, precision=prec)
        out = torch.zeros(Cn * n, dtype=torch.float64, device="cuda")
        btk.solve(cfg, opts, d[0], d[1], d[2], d[3], out)
        res[prec] = out.cpu().numpy()
    ds = torch.zeros(Cn * n, dtype=torch.float64, device="cuda")
    btk.direct_sum(dim, btk.options(kernel=kernel, exclude_self=excl), eps, d[0], d[1], d[2], d[0], d[1], d[2],
                   d[3], ds)
    dsn = ds.cpu().numpy()
    e = rel(res[btk.PREC_MIXED], res[btk.PREC_FP64])
    l64, l32 = rel(res[btk.PREC_FP64], dsn), rel(res[btk.PREC_MIXED], dsn)
    print(f"  mixed vs fp64 {e:.2e}; rel-L2 vs direct sum: fp64 {l64:.2e} mixed {l32:.2e}", flush=True)
    assert e < MIXED_TOL
    assert l32 < l64 + MIXED_TOL


def test_lists_equal_hand_derived():
    """The GPU traversal against the hand-derived lists of tests/golden/tiny_lists.json."""
    import json
    import os
    here = os.path.dirname(__file__)
    doc = json.load(open(os.path.join(here, "golden", "tiny_lists.json")))
    trees = {t["name"]: t for t in json.load(open(os.path.join(here, "golden", "tiny_trees.json")))["trees"]}
    trees.update(doc["sources"])
    for case in doc["cases"]:
        print(f"\n[hand list] {case['name']}", flush=True)
        tdef = trees[case["tree"]]
        dim, thresh = tdef["dim"], tdef["thresh"]
        cfg = btk.make_config(dim, 2, thresh, case["theta"], 0.0)

        def tree(d):
            return btk.Tree(cfg, dev(np.array(d["x"])), dev(np.array(d["y"])),
                            dev(np.array(d["z"])) if dim == 3 else None)
        tt = tree(tdef)
        ts = tree(trees[case["source"]]) if "source" in case else tt
        pairs = btk.InteractionList(cfg, tt, ts).pairs()
        got = [[int(e["target_panel"]), int(e["source_panel"]), int(e["interact_type"])] for e in pairs]
        assert got == case["list"], case["name"]
