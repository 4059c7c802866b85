"""Solve plans (btk_plan_*): the captured graph, replayed for new charge
vectors with eager work and synchronizations in between, gives the eager
solve's result every time (up to the order of FP64 atomic adds), also after
the output buffer changes (a new capture)."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("dim,kernel,prec", [(3, btk.KERNEL_POT_GRAD, btk.PREC_FP64),
                                             (3, btk.KERNEL_REF_VEL3D, btk.PREC_MIXED),
                                             (2, btk.KERNEL_POT_GRAD, btk.PREC_MIXED)])
def test_plan_replays_match_eager_solves(dim, kernel, prec):
    print(f"\n[plan] dim={dim} kernel={kernel} prec={prec}", flush=True)
    n = 30000
    if dim == 3:
        x, y, z, _ = wl.plummer_3d(n, 21, a=0.3, rmax=3.0)
    else:
        x, y, z, _ = wl.gaussian_mixture_2d(n, 21)
    d = [None if a is None else torch.from_numpy(a).cuda() for a in (x, y, z)]
    cfg = btk.make_config(dim, 6, 200, 0.7, 0.0)
    o = btk.options(kernel=kernel, exclude_self=1, precision=prec)
    C = btk.components(dim, kernel)
    plan = btk.Plan(cfg, o, d[0], d[1], d[2])
    q = torch.empty(n, dtype=torch.float64, device="cuda")
    outs = [torch.zeros(C * n, dtype=torch.float64, device="cuda") for _ in range(2)]
    g = torch.Generator(device="cuda").manual_seed(3)
    for rep in range(4):
        q.copy_(torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1)  # eager work
        out = outs[rep // 2]  # the second buffer forces a new capture
        out.zero_()
        plan.execute(q, out)
        torch.cuda.synchronize()
        ref = torch.zeros_like(out)
        btk.solve(cfg, o, d[0], d[1], d[2], q, ref)  # an eager solve between replays
        rel = float(torch.linalg.norm(out - ref) / torch.linalg.norm(ref))
        print(f"  replay {rep}: rel {rel:.2e}", flush=True)
        assert rel < 1e-13
    plan.close()


def test_plan_with_reference_options_and_bad_arguments():
    """Reference options (self term included, so eps > 0); a coincident plan
    with different target and source counts is refused."""
    print("\n[plan] arguments", flush=True)
    x, y, z, q = wl.uniform_cube(5000, 3, 4)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    cfg = btk.make_config(3, 4, 100, 0.7, 1e-3)
    with pytest.raises(btk.BtkError, match="same count"):
        h = btk.C.c_void_p()
        btk._check(btk.lib().btk_plan_create(btk.C.byref(cfg), btk.C.byref(btk.options()), 10, btk._ptr(d[0]),
                                             btk._ptr(d[1]), btk._ptr(d[2]), 20, btk._ptr(d[0]), btk._ptr(d[1]),
                                             btk._ptr(d[2]), 1, None, btk.C.byref(h)))
    plan = btk.Plan(cfg, btk.options(), d[0], d[1], d[2])
    out = torch.zeros(5000, dtype=torch.float64, device="cuda")
    ref = torch.zeros_like(out)
    plan.execute(d[3], out)
    btk.solve(cfg, btk.options(), d[0], d[1], d[2], d[3], ref)
    assert float(torch.linalg.norm(out - ref) / torch.linalg.norm(ref)) < 1e-13
    plan.close()
