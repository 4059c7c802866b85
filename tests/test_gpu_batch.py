"""Many independent problems in one forest (btk_solve_batch): every problem's
result equals its own btk_solve (FP64: 1e-12; mixed: 1e-6 of the FP64 solve),
for problems of very different sizes, including one that is a single leaf."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("dim,kernel,prec,excl,eps", [(3, btk.KERNEL_POT_GRAD, btk.PREC_FP64, 1, 0.0),
                                                      (3, btk.KERNEL_POT, btk.PREC_MIXED, 0, 1e-4),
                                                      (2, btk.KERNEL_POT_GRAD, btk.PREC_FP64, 1, 0.0)])
def test_batch_equals_individual_solves(dim, kernel, prec, excl, eps):
    print(f"\n[batch] dim={dim} kernel={kernel} prec={prec}", flush=True)
    rng = np.random.default_rng(5)
    sizes = [int(s) for s in np.exp(rng.uniform(np.log(200), np.log(20000), 24))] + [50]
    pts = []
    for i, m in enumerate(sizes):
        if dim == 3:
            x, y, z, q = wl.plummer_3d(m, 100 + i, a=0.5, rmax=5.0)
        else:
            x, y, z, q = wl.gaussian_mixture_2d(m, 100 + i)
        pts.append((x, y, z, q))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    cat = [np.concatenate([p[k] for p in pts]) if pts[0][k] is not None else None for k in range(4)]
    d = [None if a is None else torch.from_numpy(a).cuda() for a in cat]
    cfg = btk.make_config(dim, 5, 150, 0.7, eps)
    C = btk.components(dim, kernel)
    out = torch.zeros(C * off[-1], dtype=torch.float64, device="cuda")
    st = btk.solve_batch(cfg, btk.options(kernel=kernel, exclude_self=excl, precision=prec), off, d[0], d[1], d[2],
                         d[3], out)
    got = out.view(C, -1).cpu().numpy()
    worst = 0.0
    for i, (x, y, z, q) in enumerate(pts):
        m = sizes[i]
        dd = [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z, q)]
        ref = torch.zeros(C * m, dtype=torch.float64, device="cuda")
        btk.solve(cfg, btk.options(kernel=kernel, exclude_self=excl), dd[0], dd[1], dd[2], dd[3], ref)
        ref = ref.view(C, m).cpu().numpy()
        e = np.linalg.norm(got[:, off[i]:off[i + 1]] - ref) / np.linalg.norm(ref)
        worst = max(worst, e)
    print(f"  {len(sizes)} problems, {st.panel_count_source} panels; worst rel vs individual FP64 solves {worst:.2e}",
          flush=True)
    assert worst < (1e-6 if prec == btk.PREC_MIXED else 1e-12)


def test_batch_trees_are_the_individual_trees():
    """Each problem's panels in the forest are its own tree's panels (ids aside)."""
    print("\n[batch] forest structure", flush=True)
    sizes = [3000, 40, 700, 12000]
    pts = [wl.uniform_cube(m, 3, 7 + i) for i, m in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    cat = [torch.from_numpy(np.concatenate([p[k] for p in pts])).cuda() for k in range(4)]
    cfg = btk.make_config(3, 4, 100, 0.7, 0.0)
    out = torch.zeros(off[-1], dtype=torch.float64, device="cuda")
    st = btk.solve_batch(cfg, btk.options(), off, cat[0], cat[1], cat[2], cat[3], out)
    total = 0
    for i, (x, y, z, q) in enumerate(pts):
        t = btk.Tree(cfg, *[torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z)])
        total += t.panel_count
    assert st.panel_count_source == total


def test_batch_mixed_with_many_cc_pairs():
    """Mixed precision, pot+grad: a forest of problems whose boxes differ in
    size, each with enough CC pairs for the shared-kernel-matrix products
    (keyed per problem, so no key mixes box sizes); every problem within 1e-6
    of its FP64 solve."""
    print("\n[batch] mixed, large problems of different extents", flush=True)
    sizes = [60000, 150000, 90000, 120000]
    pts = []
    for i, m in enumerate(sizes):
        x, y, z, q = wl.plummer_3d(m, 300 + i, a=0.5, rmax=5.0)
        sc = 0.5 + 0.7 * i  # every problem has its own root box
        pts.append((x * sc, y * sc, z * sc, q))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    cat = [torch.from_numpy(np.concatenate([p[k] for p in pts])).cuda() for k in range(4)]
    cfg = btk.make_config(3, 6, 200, 0.7, 0.0)
    C = btk.components(3, btk.KERNEL_POT_GRAD)
    out = torch.zeros(C * off[-1], dtype=torch.float64, device="cuda")
    sb = btk.solve_batch(cfg, btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1, precision=btk.PREC_MIXED),
                         off, cat[0], cat[1], cat[2], cat[3], out)
    print(f"  forest: {int(sb.count[3])} CC pairs, {int(sb.cc_shared)} through shared kernel matrices", flush=True)
    assert sb.cc_shared > 0.3 * sb.count[3]  # keys per problem: the forest shares matrices too
    got = out.view(C, -1).cpu().numpy()
    worst = 0.0
    for i, (x, y, z, q) in enumerate(pts):
        m = sizes[i]
        dd = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z, q)]
        ref = torch.zeros(C * m, dtype=torch.float64, device="cuda")
        st = btk.solve(cfg, btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1), dd[0], dd[1], dd[2], dd[3], ref)
        ref = ref.view(C, m).cpu().numpy()
        e = np.linalg.norm(got[:, off[i]:off[i + 1]] - ref) / np.linalg.norm(ref)
        print(f"  problem {i}: {m} points, {int(st.count[3])} CC pairs, rel vs FP64 {e:.2e}", flush=True)
        worst = max(worst, e)
    assert worst < 1e-6
