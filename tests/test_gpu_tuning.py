"""Launch shapes: the autotuner installs one of its candidates, and results
do not depend on the shapes beyond the order of FP64 atomic adds; CC pairs
that share a kernel matrix give the pairwise result when run as one MFMA
product."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ITEMS = {8192, 16384, 32768, 65536, 131072}


@pytest.fixture(autouse=True)
def _restore_tuning():
    old = btk.get_tuning()
    yield
    btk.restore_tuning(old)


def _problem(dim, n, seed):
    if dim == 3:
        x, y, z, q = wl.plummer_3d(n, seed, a=0.3, rmax=3.0)
    else:
        x, y, z, q = wl.gaussian_mixture_2d(n, seed)
    return [None if a is None else torch.from_numpy(a).cuda() for a in (x, y, z, q)]


@pytest.mark.parametrize("dim,prec", [(3, btk.PREC_MIXED), (2, btk.PREC_FP64)])
def test_autotune_installs_a_candidate(dim, prec):
    print(f"\n[tuning] autotune dim={dim} prec={prec}", flush=True)
    d = _problem(dim, 30000, 3)
    cfg = btk.make_config(dim, 6, 300, 0.7, 0.0)
    o = btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1, precision=prec)
    t = btk.autotune(cfg, o, d[0], d[1], d[2], d[3])
    g = btk.get_tuning()
    assert (g.near_items, g.far_items, g.near_tpt, g.far_tpt, g.mutual, g.near_waves, g.far_waves) == \
        (t.near_items, t.far_items, t.near_tpt, t.far_tpt, t.mutual, t.near_waves, t.far_waves)
    assert t.near_items in ITEMS and t.far_items in ITEMS
    assert t.near_waves in (1, 2, 4) and t.far_waves in (1, 2, 4)
    if prec == btk.PREC_MIXED:
        assert t.near_tpt in (1, 2) and t.far_tpt in (1, 2) and t.mutual in (0, 1)
    else:
        assert t.near_tpt == 0 and t.far_tpt == 0 and t.mutual == 0


def test_autotune_multi_installs_a_candidate():
    """btk_autotune_multi picks 32, 64 or 128 charge vectors per workgroup
    and keeps every other field; the rest of the tuning is untouched."""
    print("\n[tuning] autotune_multi", flush=True)
    d = _problem(3, 20000, 5)
    Q = torch.from_numpy(np.random.default_rng(5).standard_normal((40, 20000))).cuda()
    cfg = btk.make_config(3, 6, 300, 0.7, 0.0)
    o = btk.options(kernel=btk.KERNEL_POT, exclude_self=1, precision=btk.PREC_MIXED)
    btk.set_tuning(near_items=16384, near_waves=2, cc_group=4)
    t = btk.autotune_multi(cfg, o, d[0], d[1], d[2], Q)
    g = btk.get_tuning()
    print(f"  chosen: {t.multi_vectors} vectors per workgroup", flush=True)
    assert t.multi_vectors in (32, 64, 128) and g.multi_vectors == t.multi_vectors
    assert (g.near_items, g.near_waves, g.cc_group) == (16384, 2, 4)
    with pytest.raises(btk.BtkError, match="mixed precision"):
        btk.autotune_multi(cfg, btk.options(kernel=btk.KERNEL_POT, exclude_self=1), d[0], d[1], d[2], Q)


@pytest.mark.parametrize("dim", [2, 3])
def test_results_independent_of_launch_shapes(dim):
    print(f"\n[tuning] shapes dim={dim}", flush=True)
    d = _problem(dim, 40000, 4)
    cfg = btk.make_config(dim, 6, 300, 0.7, 0.0)
    C = btk.components(dim, btk.KERNEL_POT_GRAD)
    outs = {}
    for prec in (btk.PREC_FP64, btk.PREC_MIXED):
        for shape in [(0, 0, 0, 0, 0), (8192, 8192, 1, 1, 0), (131072, 131072, 2, 2, 1), (16384, 65536, 1, 2, 1),
                      (8192, 16384, 1, 1, 0, 0, 4, 4), (32768, 8192, 2, 1, 1, 0, 2, 4)]:
            btk.set_tuning(*shape)
            out = torch.zeros(C * 40000, dtype=torch.float64, device="cuda")
            btk.solve(cfg, btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1, precision=prec), d[0], d[1],
                      d[2], d[3], out)
            outs[(prec, shape)] = out
    base = outs[(btk.PREC_FP64, (0, 0, 0, 0, 0))]
    for (prec, shape), out in outs.items():
        rel = float(torch.linalg.norm(out - base) / torch.linalg.norm(base))
        # FP64: only the order of atomic adds differs; mixed: within its own tolerance of FP64
        assert rel < (1e-13 if prec == btk.PREC_FP64 else 1e-6), (prec, shape, rel)
    fp64 = [outs[k] for k in outs if k[0] == btk.PREC_FP64]
    for out in fp64[1:]:
        assert float(torch.linalg.norm(out - fp64[0]) / torch.linalg.norm(fp64[0])) < 1e-13


@pytest.mark.parametrize("dim,kernel,leaf", [(2, btk.KERNEL_POT, 250), (2, btk.KERNEL_POT_GRAD, 250),
                                             (3, btk.KERNEL_POT, 250), (3, btk.KERNEL_REF_VEL3D, 250),
                                             (3, btk.KERNEL_POT_GRAD, 250), (3, btk.KERNEL_POT_GRAD, 2000)])
def test_mutual_pairs_match_one_way_evaluation(dim, kernel, leaf):
    """Mixed precision: evaluating mutual PP pairs once for both directions
    gives the one-way (per entry) result to FP32 rounding, for every kernel
    and both lane shapes; both are within 1e-6 of FP64.  The leaf-2000 case
    (64 uniform-cube leaves of about 1900 points) takes 8 wavefronts per
    workgroup, the largest reduction buffer; the leaf-250 cases at most 4."""
    print(f"\n[tuning] mutual dim={dim} kernel={kernel} leaf={leaf}", flush=True)
    if leaf <= 250:
        n = 40000
        d = _problem(dim, n, 5)
    else:
        n = 120000
        d = [torch.from_numpy(a).cuda() for a in wl.uniform_cube(n, 3, 5)]
    cfg = btk.make_config(dim, 5, leaf, 0.7, 0.0)
    C = btk.components(dim, kernel)

    def run(prec, **tu):
        btk.set_tuning(**tu)
        out = torch.zeros(C * n, dtype=torch.float64, device="cuda")
        btk.solve(cfg, btk.options(kernel=kernel, exclude_self=1, precision=prec), d[0], d[1], d[2], d[3], out)
        return out

    ref = run(btk.PREC_FP64)
    one_way = run(btk.PREC_MIXED)
    for tpt in (1, 2):
        mutual = run(btk.PREC_MIXED, near_tpt=tpt, mutual=1)
        rel = float(torch.linalg.norm(mutual - one_way) / torch.linalg.norm(one_way))
        assert rel < 1e-6, (tpt, rel)
        assert float(torch.linalg.norm(mutual - ref) / torch.linalg.norm(ref)) < 1e-6


@pytest.mark.parametrize("kernel,degree,learned", [(btk.KERNEL_POT_GRAD, 8, False), (btk.KERNEL_POT, 6, False),
                                                   (btk.KERNEL_POT_GRAD, 8, True)])
def test_shared_matrix_cc_matches_pairwise(kernel, degree, learned):
    """Mixed precision, 3D: the CC pairs grouped by (levels, centre offset)
    and run as MFMA products against the same list evaluated pair by pair
    (cc_group = -1), and both against the FP64 solve.  The grouped run must
    take a large share of the CC pairs; the learned case mixes degrees 4, 6
    and 8 in one list."""
    print(f"\n[tuning] shared-matrix CC kernel={kernel} degree={degree} learned={learned}", flush=True)
    n = 200000
    d = _problem(3, n, 8)
    cfg = btk.make_config(3, degree, 200, 0.7, 0.0)
    C = btk.components(3, kernel)
    old_w = None
    if learned:  # a network whose predicted error depends on the separation only
        from oracle import mlp
        try:
            old_w = btk.mac_get_weights()
        except btk.BtkError:
            pass
        w = np.zeros(mlp.N_PARAMS, np.float32)
        v = mlp.unpack(w)
        v["W1"][0, 0] = 1.0
        v["W2"][0, 0] = 1.0
        for ty in range(3):
            for k, c in enumerate((3.1, 3.7, 4.5, 5.3)):
                v["W3"][ty * 4 + k, 0] = 4.0
                v["b3"][ty * 4 + k] = -c
        btk.mac_set_weights(w)
        btk.mac_set_budget(1e-6)
    outs, stats = {}, {}
    for name, prec, cc in (("fp64", btk.PREC_FP64, 0), ("pairwise", btk.PREC_MIXED, -1),
                           ("shared", btk.PREC_MIXED, 0), ("shared2", btk.PREC_MIXED, 2)):
        btk.set_tuning(cc_group=cc)
        o = btk.options(kernel=kernel, exclude_self=1, precision=prec, mac=int(learned))
        out = torch.zeros(C * n, dtype=torch.float64, device="cuda")
        stats[name] = btk.solve(cfg, o, d[0], d[1], d[2], d[3], out)
        outs[name] = out
    if learned:
        btk.mac_set_budget(1e-7)
        if old_w is not None:
            btk.mac_set_weights(old_w)
    cc_pairs = int(stats["shared"].count[3])
    print(f"  CC pairs {cc_pairs}: shared {stats['shared'].cc_shared} (group >= 8), "
          f"{stats['shared2'].cc_shared} (group >= 2)", flush=True)
    assert stats["fp64"].cc_shared == 0 and stats["pairwise"].cc_shared == 0
    assert stats["shared"].cc_shared >= 0.3 * cc_pairs
    assert stats["shared2"].cc_shared >= stats["shared"].cc_shared
    ref = outs["fp64"]
    for name in ("pairwise", "shared", "shared2"):
        rel = float(torch.linalg.norm(outs[name] - ref) / torch.linalg.norm(ref))
        print(f"  {name}: rel vs FP64 {rel:.2e}", flush=True)
        assert rel <= 1e-6
    rel = float(torch.linalg.norm(outs["shared"] - outs["pairwise"]) / torch.linalg.norm(ref))
    print(f"  shared vs pairwise {rel:.2e}", flush=True)
    assert rel <= 5e-7


@pytest.mark.parametrize("degree,split", [(10, False), (8, True)])
def test_shared_matrix_cc_degree10_and_separate_trees(degree, split):
    """The shared-matrix CC products at the largest degree (1331 proxies per
    cluster), and with separate target and source trees (their own root
    boxes: only centre offsets on the common lattice are keyed), against the
    pairwise evaluation of the same list and the FP64 solve."""
    print(f"\n[tuning] shared-matrix CC degree={degree} split={split}", flush=True)
    n = 120000
    d = _problem(3, n, 12)
    if split:
        xs, ys, zs, qs = wl.plummer_3d(90000, 13, a=0.3, rmax=3.0)
        s = [torch.from_numpy(a).cuda() for a in (xs, ys, zs, qs)]
    cfg = btk.make_config(3, degree, 300, 0.7, 0.0)
    C = btk.components(3, btk.KERNEL_POT_GRAD)
    outs, stats = {}, {}
    for name, prec, cc in (("fp64", btk.PREC_FP64, 0), ("pairwise", btk.PREC_MIXED, -1),
                           ("shared", btk.PREC_MIXED, 2)):
        btk.set_tuning(cc_group=cc)
        o = btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=int(not split), precision=prec)
        out = torch.zeros(C * n, dtype=torch.float64, device="cuda")
        if split:
            stats[name] = btk.solve(cfg, o, d[0], d[1], d[2], s[3], out, xs=s[0], ys=s[1], zs=s[2])
        else:
            stats[name] = btk.solve(cfg, o, d[0], d[1], d[2], d[3], out)
        outs[name] = out
    print(f"  CC pairs {int(stats['shared'].count[3])}, shared {int(stats['shared'].cc_shared)}", flush=True)
    if not split:
        assert stats["shared"].cc_shared >= 0.3 * stats["shared"].count[3]
    ref = outs["fp64"]
    for name in ("pairwise", "shared"):
        rel = float(torch.linalg.norm(outs[name] - ref) / torch.linalg.norm(ref))
        print(f"  {name}: rel vs FP64 {rel:.2e}", flush=True)
        assert rel <= 1e-6
    rel = float(torch.linalg.norm(outs["shared"] - outs["pairwise"]) / torch.linalg.norm(ref))
    print(f"  shared vs pairwise {rel:.2e}", flush=True)
    assert rel <= 5e-7
