"""btk_upward_pass_3d_noalloc, the upward pass that allocates no device
memory (its scratch is a fixed array per GPU), against btk_upward_pass_3d on
the same trees and charges: bit-identical proxy weights, over two calls in a
row (the second on a smaller tree, so stale scratch contents beyond its points
must not matter), and a clear error for a tree larger than the scratch."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_noalloc_upward_equals_upward_pass():
    for n, deg, leaf, seed in ((300000, 8, 300, 1), (40000, 6, 200, 2)):
        print(f"\n[noalloc] upward pass n={n} degree={deg}", flush=True)
        x, y, z, q = (torch.from_numpy(a).cuda() for a in wl.plummer_3d(n, seed, a=0.3, rmax=5.0))
        cfg = btk.make_config(3, deg, leaf, 0.7, 0.0)
        t = btk.Tree(cfg, x, y, z)
        P = (deg + 1) ** 3
        ref = torch.zeros(t.panel_count * P, dtype=torch.float64, device="cuda")
        btk.upward_pass_3d(cfg, t, q, ref)
        # the pass writes every panel's weights: start from NaN
        got = torch.full_like(ref, float("nan"))
        stream = torch.cuda.current_stream()
        btk.upward_pass_3d_noalloc(cfg, t, q, got, stream=stream)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), float(torch.nan_to_num(got - ref, nan=np.inf).abs().max())


def test_noalloc_refuses_more_points_than_its_scratch():
    n = btk.NOALLOC_MAX_POINTS + 1
    print(f"\n[noalloc] refusal at n={n}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(3)
    x, y, z, q = (torch.rand(n, dtype=torch.float64, device="cuda", generator=g) for _ in range(4))
    cfg = btk.make_config(3, 4, 2000, 0.7, 0.0)
    t = btk.Tree(cfg, x, y, z)
    W = torch.zeros(t.panel_count * 125, dtype=torch.float64, device="cuda")
    with pytest.raises(btk.BtkError, match="BTK_NOALLOC_MAX_POINTS"):
        btk.upward_pass_3d_noalloc(cfg, t, q, W)
    del t, x, y, z, q, W
    # usable afterwards
    xs, ys, zs, qs = (torch.from_numpy(a).cuda() for a in wl.uniform_cube(5000, 3, 4))
    cfg = btk.make_config(3, 4, 100, 0.7, 0.0)
    t = btk.Tree(cfg, xs, ys, zs)
    ref = torch.zeros(t.panel_count * 125, dtype=torch.float64, device="cuda")
    got = torch.zeros_like(ref)
    btk.upward_pass_3d(cfg, t, qs, ref)
    btk.upward_pass_3d_noalloc(cfg, t, qs, got)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
