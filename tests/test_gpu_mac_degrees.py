"""Per-pair interpolation degrees of the learned MAC.  The solve runs a list
whose approximated pairs carry degrees 4, 6 and 8; the oracle runs the same
list degree by degree (each degree's own upward pass, interactions and
downward pass, summed), which pins the GPU's restriction of the charges to the
lower degrees and the extension of their potentials into the configured one.
The plan and the batched solve give the solve's result on the same list."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl
from oracle import mlp
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N, THR, DEG, THETA = 30000, 150, 8, 0.7
BUDGET = 1e-6  # log10 / 4 = -1.5


@pytest.fixture(autouse=True)
def _stepped_network():
    """Installs a network whose outputs are linear in the separation ratio
    f[0]: output 4 * sep - c_d at degree index d for every type, so that with
    the budget degree 4 serves sep < 0.4, degree 6 sep < 0.55 and degree 8
    sep < 0.75, past the theta MAC's 0.7."""
    try:
        old = btk.mac_get_weights()
    except btk.BtkError:
        old = None
    w = np.zeros(mlp.N_PARAMS, np.float32)
    v = mlp.unpack(w)
    v["W1"][0, 0] = 1.0  # h1[0] = sep
    v["W2"][0, 0] = 1.0  # h2[0] = sep
    for ty in range(3):
        for d, c in enumerate((3.1, 3.7, 4.5, 5.3)):
            v["W3"][ty * 4 + d, 0] = 4.0
            v["b3"][ty * 4 + d] = -c
    btk.mac_set_weights(w)
    btk.mac_set_budget(BUDGET)
    yield
    if old is not None:
        btk.mac_set_weights(old)
    btk.mac_set_budget(1e-7)


def _problem(seed):
    x, y, z, _ = wl.plummer_3d(N, seed, a=0.3, rmax=3.0)
    q = np.random.default_rng(seed).uniform(-1, 1, N)
    return x, y, z, q


def _oracle_by_degree(pairs, kernel, x, y, z, q):
    """The oracle on a learned list: every degree's pairs (PP with degree
    index 0) through that degree's upward pass, interactions and downward
    pass; the sum over degrees."""
    ot = po.OracleTree(3, THR, x, y, z)
    o = po.opts(kernel=kernel, exclude_self=1)
    total = None
    for di in range(DEG // 2 - 1):
        sel = pairs[(pairs["interact_type"] >> 2) == di].copy()
        sel["interact_type"] &= 3
        cfg = po.make_config(3, 4 + 2 * di, THR, THETA, 0.0)
        W = po.upward_pass(cfg, ot, q)
        near, Psi = po.fmm_interactions(cfg, o, ot, ot, q, W, sel)
        out = po.downward_pass(cfg, o, ot, Psi, near)
        total = out if total is None else total + out
    return total


@pytest.mark.parametrize("kernel", [btk.KERNEL_POT, btk.KERNEL_POT_GRAD])
def test_learned_degrees_match_oracle_degree_by_degree(kernel):
    print(f"\n[mac degrees] kernel={kernel} vs oracle", flush=True)
    x, y, z, q = _problem(5)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    cfg = btk.make_config(3, DEG, THR, THETA, 0.0)
    tree = btk.Tree(cfg, d[0], d[1], d[2])
    lst = btk.InteractionList(cfg, tree, tree, learned_mac=True)
    pairs = lst.pairs()
    ty, di = pairs["interact_type"] & 3, pairs["interact_type"] >> 2
    assert np.all(di[ty == 0] == 0) and np.all(di <= 2)
    hist = [int(np.sum((ty != 0) & (di == k))) for k in range(3)]
    print(f"  approximated pairs at degree 4/6/8: {hist}", flush=True)
    assert min(hist) > 0
    # the point side of PC and CP is a leaf (at most THR points), as in the
    # reference's types; CC may join two leaves
    cnt = tree.panels()["point_count"]
    tc, sc = cnt[pairs["target_panel"]], cnt[pairs["source_panel"]]
    assert np.all(tc[ty == 1] <= THR) and np.all(sc[ty == 2] <= THR)
    assert np.all((tc[ty == 0] <= THR) & (sc[ty == 0] <= THR))
    print(f"  CC pairs of two leaves: {int(np.sum((ty == 3) & (tc <= THR) & (sc <= THR)))}", flush=True)
    C = btk.components(3, kernel)
    out = torch.zeros(C * N, dtype=torch.float64, device="cuda")
    st = btk.solve(cfg, btk.options(kernel=kernel, exclude_self=1, mac=1), d[0], d[1], d[2], d[3], out)
    assert [int(st.count[k]) for k in range(4)] == [int(np.sum(ty == k)) for k in range(4)]
    ref = _oracle_by_degree(pairs, kernel, x, y, z, q)
    got = out.view(C, N).cpu().numpy()
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"  rel {rel:.2e}", flush=True)
    assert rel <= 1e-12
    # split_interactions sorts by the type bits and keeps the degree bits
    pp, pc, cp, cc = lst.split(cfg)
    for k, part in enumerate((pp, pc, cp, cc)):
        assert np.array_equal(part, pairs[ty == k])


@pytest.mark.parametrize("cc_group", [0, 2, -1])
def test_learned_mixed_precision_vs_oracle_degree_by_degree(cc_group):
    """Mixed precision on a learned list, with its CC pairs as shared-kernel-
    matrix MFMA products (cc_group 0: keys of >= 8 pairs, 2: >= 2) or pair by
    pair (-1), against the oracle run degree by degree in FP64: within 1e-6."""
    print(f"\n[mac degrees] mixed, cc_group={cc_group}, vs oracle", flush=True)
    old = btk.get_tuning()
    x, y, z, q = _problem(8)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    cfg = btk.make_config(3, DEG, THR, THETA, 0.0)
    tree = btk.Tree(cfg, d[0], d[1], d[2])
    pairs = btk.InteractionList(cfg, tree, tree, learned_mac=True).pairs()
    C = btk.components(3, btk.KERNEL_POT_GRAD)
    out = torch.zeros(C * N, dtype=torch.float64, device="cuda")
    btk.set_tuning(cc_group=cc_group)
    try:
        st = btk.solve(cfg, btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1, mac=1, precision=btk.PREC_MIXED),
                       d[0], d[1], d[2], d[3], out)
    finally:
        btk.restore_tuning(old)
    print(f"  CC pairs {int(st.count[3])}, as shared-matrix products {int(st.cc_shared)}", flush=True)
    if cc_group == -1:
        assert st.cc_shared == 0
    elif cc_group == 2:
        assert st.cc_shared > 0
    ref = _oracle_by_degree(pairs, btk.KERNEL_POT_GRAD, x, y, z, q)
    rel = np.linalg.norm(out.view(C, N).cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"  rel vs oracle {rel:.2e}", flush=True)
    assert rel <= 1e-6


def test_cost_weights_steer_the_choice():
    """btk_mac_set_costs: with a CC evaluation priced at a tenth of the others
    the learned list takes more CC and fewer PC/CP pairs than with equal
    prices, and CC pairs with a leaf on one side.  At a hundred times no CC
    pair has a leaf side: this network predicts one error for every type, so
    PC or CP (at most THR * P evaluations) is within the budget whenever CC
    is, and cheaper.  The weighted list's mixed solve (CC as MFMA products)
    matches the oracle run degree by degree within 1e-6."""
    print("\n[mac degrees] cost weights", flush=True)
    x, y, z, q = _problem(9)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    cfg = btk.make_config(3, DEG, THR, THETA, 0.0)
    tree = btk.Tree(cfg, d[0], d[1], d[2])
    counts = {}
    pairs = {}
    try:
        for cc in (1.0, 0.1, 100.0):
            btk.mac_set_costs(1.0, 1.0, 1.0, cc)
            p = btk.InteractionList(cfg, tree, tree, learned_mac=True).pairs()
            ty = p["interact_type"] & 3
            counts[cc] = [int(np.sum(ty == k)) for k in range(4)]
            pairs[cc] = p
            print(f"  CC cost {cc}: PP/PC/CP/CC pairs {counts[cc]}", flush=True)
        assert counts[0.1][3] > counts[1.0][3]
        assert counts[0.1][1] + counts[0.1][2] < counts[1.0][1] + counts[1.0][2]
        cnt = tree.panels()["point_count"]

        def leaf_side_cc(p):
            cc = (p["interact_type"] & 3) == 3
            return int(np.sum(cc & ((cnt[p["target_panel"]] <= THR) | (cnt[p["source_panel"]] <= THR))))

        print(f"  CC pairs with a leaf side: cost 0.1 {leaf_side_cc(pairs[0.1])}, "
              f"cost 100 {leaf_side_cc(pairs[100.0])}", flush=True)
        assert leaf_side_cc(pairs[0.1]) > 0 and leaf_side_cc(pairs[100.0]) == 0
        btk.mac_set_costs(1.0, 1.0, 1.0, 0.1)
        C = btk.components(3, btk.KERNEL_POT_GRAD)
        out = torch.zeros(C * N, dtype=torch.float64, device="cuda")
        st = btk.solve(cfg, btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1, mac=1, precision=btk.PREC_MIXED),
                       d[0], d[1], d[2], d[3], out)
        assert [int(st.count[k]) for k in range(4)] == counts[0.1]
    finally:
        btk.mac_set_costs()
    ref = _oracle_by_degree(pairs[0.1], btk.KERNEL_POT_GRAD, x, y, z, q)
    rel = np.linalg.norm(out.view(C, N).cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"  CC pairs as shared-matrix products {int(st.cc_shared)}; rel vs oracle {rel:.2e}", flush=True)
    assert rel <= 1e-6


def test_learned_plan_and_batch_match_the_solve():
    print("\n[mac degrees] plan and batch vs solve", flush=True)
    x, y, z, q = _problem(6)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    cfg = btk.make_config(3, DEG, THR, THETA, 0.0)
    o = btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1, mac=1)
    C = btk.components(3, btk.KERNEL_POT_GRAD)
    ref = torch.zeros(C * N, dtype=torch.float64, device="cuda")
    btk.solve(cfg, o, d[0], d[1], d[2], d[3], ref)
    plan = btk.Plan(cfg, o, d[0], d[1], d[2])
    for rep in range(2):
        out = torch.zeros_like(ref)
        plan.execute(d[3], out)
        torch.cuda.synchronize()
        rel = float(torch.linalg.norm(out - ref) / torch.linalg.norm(ref))
        print(f"  plan replay {rep}: rel {rel:.2e}", flush=True)
        assert rel < 1e-13
    plan.close()
    # two copies of the problem in one forest
    cat = [torch.cat([a, a]) for a in d]
    out = torch.zeros(C * 2 * N, dtype=torch.float64, device="cuda")
    btk.solve_batch(cfg, o, [0, N, 2 * N], cat[0], cat[1], cat[2], cat[3], out)
    out = out.view(C, 2 * N)
    r = ref.view(C, N)
    for p in range(2):
        rel = float(torch.linalg.norm(out[:, p * N:(p + 1) * N] - r) / torch.linalg.norm(r))
        print(f"  batch problem {p}: rel {rel:.2e}", flush=True)
        assert rel < 1e-12


def test_learned_list_is_refused_by_the_fixed_degree_entry_points():
    """btk_fmm_interactions_3d takes arrays of one degree: a learned list
    mixing degrees is an error, not an out-of-bounds read."""
    print("\n[mac degrees] fixed-degree entry point refuses a learned list", flush=True)
    x, y, z, q = _problem(7)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    cfg = btk.make_config(3, DEG, THR, THETA, 0.0)
    tree = btk.Tree(cfg, d[0], d[1], d[2])
    lst = btk.InteractionList(cfg, tree, tree, learned_mac=True)
    P = (DEG + 1) ** 3
    W = torch.zeros(tree.panel_count * P, dtype=torch.float64, device="cuda")
    Psi = torch.zeros_like(W)
    out = torch.zeros(N, dtype=torch.float64, device="cuda")
    with pytest.raises(btk.BtkError, match="learned-MAC list"):
        btk.poisson_fmm_interactions_3d(cfg, tree, tree, lst, d[3], W, out, Psi, opts=btk.options(exclude_self=1))
