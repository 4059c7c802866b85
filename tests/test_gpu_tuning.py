"""Launch shapes: the autotuner installs one of its candidates, and results
do not depend on the shapes beyond the order of FP64 atomic adds."""
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
    btk.set_tuning(old.near_items, old.far_items, old.near_tpt, old.far_tpt)


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
    assert (g.near_items, g.far_items, g.near_tpt, g.far_tpt) == (t.near_items, t.far_items, t.near_tpt, t.far_tpt)
    assert t.near_items in ITEMS and t.far_items in ITEMS
    if prec == btk.PREC_MIXED:
        assert t.near_tpt in (1, 2) and t.far_tpt in (1, 2)
    else:
        assert t.near_tpt == 0 and t.far_tpt == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_results_independent_of_launch_shapes(dim):
    print(f"\n[tuning] shapes dim={dim}", flush=True)
    d = _problem(dim, 40000, 4)
    cfg = btk.make_config(dim, 6, 300, 0.7, 0.0)
    C = btk.components(dim, btk.KERNEL_POT_GRAD)
    outs = {}
    for prec in (btk.PREC_FP64, btk.PREC_MIXED):
        for shape in [(0, 0, 0, 0), (8192, 8192, 1, 1), (131072, 131072, 2, 2), (16384, 65536, 1, 2)]:
            btk.set_tuning(*shape)
            out = torch.zeros(C * 40000, dtype=torch.float64, device="cuda")
            btk.solve(cfg, btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1, precision=prec), d[0], d[1],
                      d[2], d[3], out)
            outs[(prec, shape)] = out
    base = outs[(btk.PREC_FP64, (0, 0, 0, 0))]
    for (prec, shape), out in outs.items():
        rel = float(torch.linalg.norm(out - base) / torch.linalg.norm(base))
        # FP64: only the order of atomic adds differs; mixed: within its own tolerance of FP64
        assert rel < (1e-13 if prec == btk.PREC_FP64 else 1e-6), (prec, shape, rel)
    mixed = [outs[k] for k in outs if k[0] == btk.PREC_MIXED]
    for out in mixed[1:]:
        # the mixed path rounds identically under every shape; only FP64 adds are reordered
        assert float(torch.linalg.norm(out - mixed[0]) / torch.linalg.norm(mixed[0])) < 1e-13
