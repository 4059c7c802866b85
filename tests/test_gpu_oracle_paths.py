"""The paths the headline numbers run on, against the oracle itself (the FP64
CPU restatement of poisson_fmm_impl.hpp:211-442 and the rest of the solve):

- mixed precision (FP32 kernel evaluations, FP64 accumulation), with the
  launch shapes the autotuner chooses from: one or two targets per lane,
  mutual PP pairs, CC pairs evaluated pairwise or as shared-kernel-matrix
  MFMA products;
- the sphere + Plummer potential+gradient problem at degrees 8 and 10;
- many charge vectors on MFMA (btk_solve_multi), vector by vector;
- a forest of uneven Plummer problems (btk_solve_batch), problem by problem.

FP64 results must be within 1e-12 rel-L2 of the oracle, mixed-precision ones
within 1e-6 (the metric's own bound).
"""
import functools

import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FP64_TOL = 1e-12
MIXED_TOL = 1e-6


@pytest.fixture(autouse=True)
def _restore_tuning():
    old = btk.get_tuning()
    yield
    btk.restore_tuning(old)


def points(kind, n, seed):
    if kind == "cube3":
        return wl.uniform_cube(n, 3, seed)
    if kind == "gauss2":
        return wl.gaussian_mixture_2d(n, seed)
    if kind == "sphere3":
        return wl.sphere_plummer_3d(n, seed)
    if kind == "plummer3":
        return wl.plummer_3d(n, seed, a=0.3, rmax=3.0)
    raise ValueError(kind)


@functools.lru_cache(maxsize=8)
def oracle_solve(kind, n, seed, dim, deg, leaf, theta, eps, kernel, excl):
    O.set_threads(16)
    x, y, z, q = points(kind, n, seed)
    out, _, cnt = O.solve(O.make_config(dim, deg, leaf, theta, eps), O.opts(kernel=kernel, exclude_self=excl),
                          x, y, z, x, y, z, q)
    return out, [int(c) for c in cnt]


def rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [
    # kind, n, dim, degree, leaf, theta, eps, kernel, exclude_self
    ("sphere3", 24000, 3, 8, 400, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1),   # sphere + Plummer, degree 8
    ("sphere3", 20000, 3, 10, 600, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1),  # sphere + Plummer, degree 10
    ("plummer3", 30000, 3, 8, 300, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1),  # Plummer a = 0.3
    ("cube3", 30000, 3, 6, 120, 0.7, 1e-4, btk.KERNEL_POT, 0),         # many CC pairs
    ("cube3", 30000, 3, 6, 120, 0.7, 0.0, btk.KERNEL_POT_GRAD, 1),     # many CC pairs, gradient
    ("sphere3", 16000, 3, 6, 250, 0.7, 1e-4, btk.KERNEL_REF_VEL3D, 0),
    ("gauss2", 60000, 2, 6, 500, 0.6, 0.0, btk.KERNEL_POT_GRAD, 1),    # 2D Gaussian mixture
    ("gauss2", 40000, 2, 7, 300, 0.6, 1e-4, btk.KERNEL_POT, 0),
]

# launch shapes (btk_set_tuning keywords) the mixed path is run with
SHAPES = [
    ("default", {}),
    ("two_targets_per_lane", {"near_tpt": 2, "far_tpt": 2}),
    ("mutual_pairs", {"near_tpt": 1, "mutual": 1}),
    ("cc_pairwise", {"cc_group": -1}),
    ("cc_shared_from_2", {"cc_group": 2}),
    ("four_waves_per_tile", {"near_waves": 4, "far_waves": 4}),
    ("two_waves_two_targets_per_lane", {"near_waves": 2, "far_waves": 2, "near_tpt": 2, "far_tpt": 2}),
]


@pytest.mark.parametrize("kind,n,dim,deg,leaf,theta,eps,kernel,excl", CASES)
def test_solves_vs_oracle(kind, n, dim, deg, leaf, theta, eps, kernel, excl):
    print(f"\n[oracle paths] {kind} n={n} dim={dim} deg={deg} leaf={leaf} kernel={kernel}", flush=True)
    seed = 500 + n + deg
    ref, cnt = oracle_solve(kind, n, seed, dim, deg, leaf, theta, eps, kernel, excl)
    x, y, z, q = points(kind, n, seed)
    d = [None if a is None else torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    cfg = btk.make_config(dim, deg, leaf, theta, eps)
    C = btk.components(dim, kernel)

    def run(prec, **shape):
        btk.set_tuning()
        btk.set_tuning(**shape)
        out = torch.zeros(C * n, dtype=torch.float64, device="cuda")
        st = btk.solve(cfg, btk.options(kernel=kernel, exclude_self=excl, precision=prec), d[0], d[1], d[2], d[3],
                       out)
        return out.cpu().numpy(), st

    got, st = run(btk.PREC_FP64)
    assert [int(st.count[k]) for k in range(4)] == cnt
    e = rel(got, ref)
    print(f"  pairs {cnt}; FP64 vs oracle {e:.2e}", flush=True)
    assert e < FP64_TOL
    got, st = run(btk.PREC_FP64, near_waves=4, far_waves=2)
    e = rel(got, ref)
    print(f"  FP64, 4 and 2 wavefronts per tile, vs oracle {e:.2e}", flush=True)
    assert e < FP64_TOL
    for name, shape in SHAPES:
        if dim == 2 and "cc_group" in shape:
            continue  # shared kernel matrices: 3D only
        got, st = run(btk.PREC_MIXED, **shape)
        e = rel(got, ref)
        print(f"  mixed [{name}] vs oracle {e:.2e} (CC pairs shared: {int(st.cc_shared)} of {cnt[3]})", flush=True)
        assert e < MIXED_TOL, name
        if name == "cc_shared_from_2" and cnt[3] >= 200 and kernel != btk.KERNEL_REF_VEL3D:
            assert st.cc_shared > 0.3 * cnt[3]


@pytest.mark.parametrize("se,ce", [(-16, -30), (16, 30), (-16, 30), (16, -30)])
def test_shared_cc_scale_invariance(se, ce):
    """The shared-matrix CC products carry their operands as split fp16 with
    power-of-two scalings chosen from the boxes' gap and each cluster's
    largest weight.  Coordinates scaled by 2^se and charges by 2^ce (exact in
    FP64: the same tree and lists) must give the oracle's solution scaled by
    2^(ce-se) (potential) and 2^(ce-2se) (gradient), each component as close
    as at scale 1."""
    kind, n, dim, deg, leaf, theta, eps, kernel, excl = CASES[4]
    print(f"\n[oracle paths] shared CC at coordinates x 2^{se}, charges x 2^{ce}: {kind} n={n}", flush=True)
    seed = 500 + n + deg
    ref, cnt = oracle_solve(kind, n, seed, dim, deg, leaf, theta, eps, kernel, excl)
    C = btk.components(dim, kernel)
    ref = np.asarray(ref).reshape(C, n)
    x, y, z, q = points(kind, n, seed)
    cfg = btk.make_config(dim, deg, leaf, theta, eps)
    btk.set_tuning(cc_group=2)

    def run(s, c):
        d = [torch.from_numpy(np.ascontiguousarray(a * s)).cuda() for a in (x, y, z)]
        out = torch.zeros(C * n, dtype=torch.float64, device="cuda")
        st = btk.solve(cfg, btk.options(kernel=kernel, exclude_self=excl, precision=btk.PREC_MIXED), d[0], d[1],
                       d[2], torch.from_numpy(q * c).cuda(), out)
        assert [int(st.count[k]) for k in range(4)] == cnt
        assert st.cc_shared > 0.3 * cnt[3]
        return out.view(C, n).cpu().numpy()

    base = run(1.0, 1.0)
    s, c = 2.0 ** se, 2.0 ** ce
    got = run(s, c)
    want = ref * np.array([c / s] + [c / (s * s)] * (C - 1))[:, None]
    for k in range(C):
        e0, e = rel(base[k], ref[k]), rel(got[k], want[k])
        print(f"  component {k}: vs oracle {e:.2e} (scale 1: {e0:.2e})", flush=True)
        assert e < max(2.0 * e0, 1e-9), k
    assert rel(got, want) < MIXED_TOL


MULTI = [
    # kind, n, dim, degree, leaf, theta, eps, kernel, exclude_self, vectors
    ("cube3", 20000, 3, 6, 300, 0.7, 0.0, btk.KERNEL_POT, 1, 64),
    ("sphere3", 16000, 3, 6, 300, 0.7, 1e-4, btk.KERNEL_POT_GRAD, 0, 20),
    ("gauss2", 30000, 2, 6, 400, 0.6, 0.0, btk.KERNEL_POT_GRAD, 1, 64),
]


@pytest.mark.parametrize("kind,n,dim,deg,leaf,theta,eps,kernel,excl,nrhs", MULTI)
def test_many_vectors_vs_oracle(kind, n, dim, deg, leaf, theta, eps, kernel, excl, nrhs):
    """btk_solve_multi in mixed precision (FP32 MFMA products for >= 16
    vectors), with 32, 64 and 128 charge vectors per workgroup: sampled
    vectors, each against the oracle's FP64 solve of that vector alone."""
    print(f"\n[oracle paths] many vectors {kind} n={n} dim={dim} kernel={kernel} vectors={nrhs}", flush=True)
    O.set_threads(16)
    x, y, z, _ = points(kind, n, 40 + nrhs)
    Q = np.random.default_rng(nrhs).standard_normal((nrhs, n))
    cfg = btk.make_config(dim, deg, leaf, theta, eps)
    C = btk.components(dim, kernel)
    d = [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z)]
    got = {}
    for mv in (32, 64, 128):
        btk.set_tuning(multi_vectors=mv)
        out = torch.zeros(nrhs * C * n, dtype=torch.float64, device="cuda")
        btk.solve_multi(cfg, btk.options(kernel=kernel, exclude_self=excl, precision=btk.PREC_MIXED), d[0], d[1],
                        d[2], torch.from_numpy(Q).cuda(), out)
        got[mv] = out.view(nrhs, C, n).cpu().numpy()
    ocfg = O.make_config(dim, deg, leaf, theta, eps)
    for r in (0, nrhs // 2, nrhs - 1):
        ref, _, _ = O.solve(ocfg, O.opts(kernel=kernel, exclude_self=excl), x, y, z, x, y, z, Q[r])
        for mv in (32, 64, 128):
            e = rel(got[mv][r], ref)
            print(f"  vector {r}, {mv} vectors per workgroup: vs oracle {e:.2e}", flush=True)
            assert e < MIXED_TOL, mv
    # a column's FP32 tile products are the same MFMA steps in every
    # partition of the vectors; only the FP64 adds of the tiles (and of split
    # items) may run in another order
    assert rel(got[32], got[64]) < 1e-12 and rel(got[128], got[64]) < 1e-12


@pytest.mark.parametrize("kernel,prec,deg", [(btk.KERNEL_POT_GRAD, btk.PREC_FP64, 6),
                                             (btk.KERNEL_POT_GRAD, btk.PREC_MIXED, 8),
                                             (btk.KERNEL_POT, btk.PREC_MIXED, 6)])
def test_forest_vs_oracle(kernel, prec, deg):
    """24 Plummer problems of 300 to 20,000 points (log-uniform) in one forest
    solve; every problem against the oracle's solve of that problem alone."""
    print(f"\n[oracle paths] forest kernel={kernel} prec={prec} deg={deg}", flush=True)
    O.set_threads(16)
    rng = np.random.default_rng(17)
    sizes = [int(s) for s in np.exp(rng.uniform(np.log(300), np.log(20000), 24))]
    pts = [wl.plummer_3d(m, 700 + i, a=0.5, rmax=5.0) for i, m in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    cat = [torch.from_numpy(np.concatenate([p[k] for p in pts])).cuda() for k in range(4)]
    leaf = 150
    cfg = btk.make_config(3, deg, leaf, 0.7, 0.0)
    C = btk.components(3, kernel)
    out = torch.zeros(C * int(off[-1]), dtype=torch.float64, device="cuda")
    btk.solve_batch(cfg, btk.options(kernel=kernel, exclude_self=1, precision=prec), off, cat[0], cat[1], cat[2],
                    cat[3], out)
    got = out.view(C, -1).cpu().numpy()
    ocfg = O.make_config(3, deg, leaf, 0.7, 0.0)
    worst = 0.0
    for i, (x, y, z, q) in enumerate(pts):
        ref, _, _ = O.solve(ocfg, O.opts(kernel=kernel, exclude_self=1), x, y, z, x, y, z, q)
        worst = max(worst, rel(got[:, off[i]:off[i + 1]], ref))
    print(f"  {len(sizes)} problems ({sum(sizes)} points): worst rel vs oracle {worst:.2e}", flush=True)
    assert worst < (MIXED_TOL if prec == btk.PREC_MIXED else FP64_TOL)
