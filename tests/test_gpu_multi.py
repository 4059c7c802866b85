"""Many charge vectors over one geometry (btk_solve_multi): the FP32 MFMA
products (nrhs >= 16, mixed precision) and the per-vector path against
independent FP64 single-vector solves of every vector."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MULTI_TOL = 1e-6  # rel-L2 distance of each vector's result to its FP64 single-vector solve

CASES = [
    # generator, dim, n, degree, thresh, theta, eps, kernel, excl, nrhs, precision
    ("cube3", 3, 30000, 6, 400, 0.7, 0.0, btk.KERNEL_POT, 1, 64, btk.PREC_MIXED),
    ("sphere3", 3, 20000, 6, 300, 0.7, 1e-4, btk.KERNEL_POT_GRAD, 0, 20, btk.PREC_MIXED),
    ("gauss2", 2, 40000, 6, 500, 0.6, 0.0, btk.KERNEL_POT_GRAD, 1, 64, btk.PREC_MIXED),
    ("cube2", 2, 20000, 5, 300, 0.6, 1e-4, btk.KERNEL_POT, 0, 70, btk.PREC_MIXED),
    ("cube3", 3, 10000, 5, 300, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1, 8, btk.PREC_FP64),
]


def gen(kind, n, seed):
    if kind == "cube3":
        return wl.uniform_cube(n, 3, seed)
    if kind == "cube2":
        return wl.uniform_cube(n, 2, seed)
    if kind == "gauss2":
        return wl.gaussian_mixture_2d(n, seed)
    return wl.sphere_plummer_3d(n, seed)


@pytest.mark.parametrize("kind,dim,n,deg,thresh,theta,eps,kernel,excl,nrhs,prec", CASES)
def test_multi_matches_single_vector_solves(kind, dim, n, deg, thresh, theta, eps, kernel, excl, nrhs, prec):
    print(f"\n[multi] {kind} dim={dim} n={n} deg={deg} kernel={kernel} nrhs={nrhs} prec={prec}", flush=True)
    x, y, z, _ = gen(kind, n, 900 + nrhs)
    rng = np.random.default_rng(nrhs)
    Q = rng.standard_normal((nrhs, n))
    cfg = btk.make_config(dim, deg, thresh, theta, eps)
    C = btk.components(dim, kernel)
    d = [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z)]
    dq = torch.from_numpy(Q).cuda()
    out = torch.zeros(nrhs * C * n, dtype=torch.float64, device="cuda")
    btk.solve_multi(cfg, btk.options(kernel=kernel, exclude_self=excl, precision=prec), d[0], d[1], d[2], dq, out)
    got = out.view(nrhs, C, n).cpu().numpy()
    worst = 0.0
    for r in list(range(0, nrhs, max(1, nrhs // 6))) + [nrhs - 1]:
        ref = torch.zeros(C * n, dtype=torch.float64, device="cuda")
        btk.solve(cfg, btk.options(kernel=kernel, exclude_self=excl), d[0], d[1], d[2], dq[r].contiguous(), ref)
        ref = ref.view(C, n).cpu().numpy()
        e = np.linalg.norm(got[r] - ref) / np.linalg.norm(ref)
        worst = max(worst, e)
    print(f"  worst rel-L2 vs FP64 single-vector solves {worst:.2e}", flush=True)
    assert worst < (MULTI_TOL if prec == btk.PREC_MIXED else 1e-12)
