"""The persistent work-queue schedule (btk_options.schedule =
BTK_SCHEDULE_QUEUE: every PP/PC/CP/CC work item in one launch, biggest first,
taken by atomic ticket) and the solve plan of a fixed forest replayed as one
HIP graph (btk_plan_create_batch).

The queue must give the group launches' result (FP64: up to the order of
atomic adds), single solves and forests must match the oracle (FP64 1e-12,
mixed 1e-6), and graph replays, with eager work and synchronizations between
them, must each match the eager forest solve."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("dim,kernel,prec", [(3, btk.KERNEL_POT_GRAD, btk.PREC_MIXED), (3, btk.KERNEL_POT, btk.PREC_FP64),
                                             (2, btk.KERNEL_POT_GRAD, btk.PREC_MIXED),
                                             (3, btk.KERNEL_REF_VEL3D, btk.PREC_MIXED)])
def test_queue_matches_group_launches_and_oracle(dim, kernel, prec):
    print(f"\n[queue] dim={dim} kernel={kernel} prec={prec}", flush=True)
    n = 30000
    if dim == 3:
        x, y, z, q = wl.plummer_3d(n, 41, a=0.3, rmax=3.0)
    else:
        x, y, z, q = wl.gaussian_mixture_2d(n, 41)
    d = [None if a is None else torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    cfg = btk.make_config(dim, 6, 250, 0.7, 0.0)
    C = btk.components(dim, kernel)
    outs = {}
    for sched in (btk.SCHEDULE_GROUPS, btk.SCHEDULE_QUEUE):
        out = torch.zeros(C * n, dtype=torch.float64, device="cuda")
        st = btk.solve(cfg, btk.options(kernel=kernel, exclude_self=1, precision=prec, schedule=sched), d[0], d[1],
                       d[2], d[3], out)
        outs[sched] = out.cpu().numpy()
        print(f"  schedule {sched}: near {st.t_kernel_near * 1e3:.2f} ms, far {st.t_kernel_far * 1e3:.2f} ms", flush=True)
    e = rel(outs[btk.SCHEDULE_QUEUE], outs[btk.SCHEDULE_GROUPS])
    print(f"  queue vs group launches {e:.2e}", flush=True)
    assert e < (1e-13 if prec == btk.PREC_FP64 else 1e-6)
    O.set_threads(16)
    ref, _, _ = O.solve(O.make_config(dim, 6, 250, 0.7, 0.0), O.opts(kernel=kernel, exclude_self=1), x, y, z, x, y,
                        z, q)
    e = rel(outs[btk.SCHEDULE_QUEUE], ref)
    print(f"  queue vs oracle {e:.2e}", flush=True)
    assert e < (1e-12 if prec == btk.PREC_FP64 else 1e-6)


def test_queue_with_learned_degrees():
    """A learned-MAC list (degrees 4, 6 and 8 in one solve): near groups of
    several degrees write the same targets from one launch."""
    print("\n[queue] learned degrees", flush=True)
    from oracle import mlp
    old = btk.mac_get_weights() if _has_weights() else None
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
    try:
        n = 40000
        x, y, z, q = wl.plummer_3d(n, 43, a=0.3, rmax=3.0)
        d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
        cfg = btk.make_config(3, 8, 150, 0.7, 0.0)
        C = btk.components(3, btk.KERNEL_POT_GRAD)
        res = {}
        for prec in (btk.PREC_FP64, btk.PREC_MIXED):
            for sched in (btk.SCHEDULE_GROUPS, btk.SCHEDULE_QUEUE):
                out = torch.zeros(C * n, dtype=torch.float64, device="cuda")
                btk.solve(cfg, btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1, precision=prec, mac=1,
                                           schedule=sched), d[0], d[1], d[2], d[3], out)
                res[(prec, sched)] = out.cpu().numpy()
        e64 = rel(res[(btk.PREC_FP64, btk.SCHEDULE_QUEUE)], res[(btk.PREC_FP64, btk.SCHEDULE_GROUPS)])
        e32 = rel(res[(btk.PREC_MIXED, btk.SCHEDULE_QUEUE)], res[(btk.PREC_FP64, btk.SCHEDULE_GROUPS)])
        print(f"  FP64 queue vs groups {e64:.2e}; mixed queue vs FP64 {e32:.2e}", flush=True)
        assert e64 < 1e-13 and e32 < 1e-6
    finally:
        btk.mac_set_budget(1e-7)
        if old is not None:
            btk.mac_set_weights(old)


def _has_weights():
    try:
        btk.mac_get_weights()
        return True
    except btk.BtkError:
        return False


def _forest(nprob, seed, lo=300, hi=20000):
    rng = np.random.default_rng(seed)
    sizes = [int(s) for s in np.exp(rng.uniform(np.log(lo), np.log(hi), nprob))]
    pts = [wl.plummer_3d(m, seed * 100 + i, a=0.5, rmax=5.0) for i, m in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    cat = [torch.from_numpy(np.concatenate([p[k] for p in pts])).cuda() for k in range(4)]
    return sizes, pts, off, cat


@pytest.mark.parametrize("prec,kernel", [(btk.PREC_FP64, btk.KERNEL_POT_GRAD), (btk.PREC_MIXED, btk.KERNEL_POT)])
def test_forest_plan_graph_with_queue_vs_oracle(prec, kernel):
    print(f"\n[queue] forest plan, prec={prec} kernel={kernel}", flush=True)
    sizes, pts, off, cat = _forest(24, 7)
    cfg = btk.make_config(3, 6, 150, 0.7, 0.0)
    C = btk.components(3, kernel)
    n = int(off[-1])
    o = btk.options(kernel=kernel, exclude_self=1, precision=prec, schedule=btk.SCHEDULE_QUEUE)
    eager = torch.zeros(C * n, dtype=torch.float64, device="cuda")
    btk.solve_batch(cfg, o, off, cat[0], cat[1], cat[2], cat[3], eager)
    plan = btk.Plan.batch(cfg, o, off, cat[0], cat[1], cat[2])
    tol_replay = 1e-13 if prec == btk.PREC_FP64 else 1e-6
    for rep in range(3):
        out = torch.zeros(C * n, dtype=torch.float64, device="cuda")
        plan.execute(cat[3], out)
        torch.cuda.synchronize()
        e = float(torch.linalg.norm(out - eager) / torch.linalg.norm(eager))
        print(f"  replay {rep}: vs eager forest solve {e:.2e}", flush=True)
        assert e < tol_replay
        junk = torch.randn(1 << 20, device="cuda") * 2  # eager work between replays
        torch.cuda.synchronize()
        del junk
    # a second charge vector through the same plan (new capture)
    q2 = cat[3] * -0.5 + 0.25
    out2 = torch.zeros(C * n, dtype=torch.float64, device="cuda")
    plan.execute(q2, out2)
    torch.cuda.synchronize()
    plan.close()
    got = out.view(C, n).cpu().numpy()
    got2 = out2.view(C, n).cpu().numpy()
    O.set_threads(16)
    ocfg = O.make_config(3, 6, 150, 0.7, 0.0)
    worst = worst2 = 0.0
    q2h = q2.cpu().numpy()
    for i, (x, y, z, q) in enumerate(pts):
        ref, _, _ = O.solve(ocfg, O.opts(kernel=kernel, exclude_self=1), x, y, z, x, y, z, q)
        worst = max(worst, rel(got[:, off[i]:off[i + 1]], ref))
        if i % 6 == 0:
            ref2, _, _ = O.solve(ocfg, O.opts(kernel=kernel, exclude_self=1), x, y, z, x, y, z,
                                 q2h[off[i]:off[i + 1]])
            worst2 = max(worst2, rel(got2[:, off[i]:off[i + 1]], ref2))
    print(f"  {len(sizes)} problems: worst rel vs oracle {worst:.2e}, second vector {worst2:.2e}", flush=True)
    tol = 1e-12 if prec == btk.PREC_FP64 else 1e-6
    assert worst < tol and worst2 < tol
