"""Pins the oracle's upward pass, interactions, downward pass and direct sum
without the reference binary:

- upward (upward_pass_impl.hpp): every panel's proxy weights reproduce the
  moments of degree <= n of the charges in its subtree (anterpolation and
  child->parent are exact on polynomials), including the total charge;
- interactions (poisson_fmm_impl.hpp, poisson_vel_fmm_impl.hpp): a list of one
  pair equals the textbook sum over that pair's points / proxy points,
  evaluated here with numpy from the kernel's closed form;
- kernels: sign conventions at known distances (F2), the gradient against
  central differences of the potential, the reference "vel" kernel's closed
  form (F4);
- downward (downward_pass_impl.hpp): a field of degree <= n given on the root's
  target proxies reaches every point exactly (parent->child and leaf->point
  interpolate polynomials exactly);
- whole path: the error against the direct sum falls as the degree rises;
- degree change: the lower-degree upward and downward passes from the
  higher-degree ones through the 1-D basis matrix between the two degrees.
"""
import numpy as np
import pytest

from barytreek_amd import workloads as wl
from oracle import pyoracle as O

PI = np.pi


def grid(lo, hi, n, dim):
    c = [np.cos(PI / n * np.arange(n + 1)) * 0.5 * (hi[d] - lo[d]) + 0.5 * (hi[d] + lo[d]) for d in range(dim)]
    g = np.meshgrid(*c, indexing="ij")
    return [a.ravel() for a in g]


def panels(t):
    p = t.t.panels
    lo = np.array([[p[i].lo[d] for d in range(3)] for i in range(t.panel_count)])
    hi = np.array([[p[i].hi[d] for d in range(3)] for i in range(t.panel_count)])
    child = np.array([[p[i].child[c] for c in range(8)] for i in range(t.panel_count)])
    leaf = np.array([p[i].is_leaf for i in range(t.panel_count)], bool)
    return lo, hi, child, leaf


def subtree(t, p, child, leaf, nc):
    if leaf[p]:
        return t.leaf_points[t.leaf_start[p]:t.leaf_start[p + 1]]
    return np.concatenate([subtree(t, c, child, leaf, nc) for c in child[p, :nc]])


@pytest.mark.parametrize("dim,n,deg,thresh", [(3, 4000, 5, 80), (2, 6000, 7, 60), (3, 2000, 8, 150)])
def test_upward_reproduces_moments(dim, n, deg, thresh):
    x, y, z, q = wl.uniform_cube(n, dim, 8)
    t = O.OracleTree(dim, thresh, x, y, z)
    cfg = O.make_config(dim, deg, thresh, 0.7, 0.0)
    W = O.upward_pass(cfg, t, q)
    lo, hi, child, leaf = panels(t)
    co = [x, y, z]
    rng = np.random.default_rng(0)
    for p in rng.choice(t.panel_count, 25, replace=False):
        pts = subtree(t, p, child, leaf, 2 ** dim)
        g = grid(lo[p], hi[p], deg, dim)
        assert abs(W[p].sum() - q[pts].sum()) < 1e-11 * (1 + np.abs(q[pts]).sum())
        for _ in range(3):
            e = rng.integers(0, deg + 1, dim)  # monomial of degree <= deg per axis
            mono_pts = np.prod([co[d][pts] ** e[d] for d in range(dim)], axis=0)
            mono_grid = np.prod([g[d] ** e[d] for d in range(dim)], axis=0)
            ref = (q[pts] * mono_pts).sum()
            assert abs(W[p] @ mono_grid - ref) < 1e-10 * (1 + np.abs(q[pts]).sum())


def G(dim, kernel, conv, eps, t, s, q):
    """Closed forms (poisson_fmm_impl.hpp:38,47,241,252; poisson_vel_fmm_impl.hpp:42,53)."""
    d = t[:, None, :] - s[None, :, :]
    r2 = (d ** 2).sum(-1) + eps * eps
    sign = -1.0 if conv else 1.0
    if dim == 2:
        g = sign / (4 * PI)
        out = [(g * np.log(r2) * q).sum(1)]
        if kernel == O.KERNEL_POT_GRAD:
            out += [(2 * g * d[..., k] / r2 * q).sum(1) for k in range(2)]
        return np.array(out)
    g = -sign / (4 * PI)
    if kernel == O.KERNEL_POT:
        return np.array([(g / np.sqrt(r2) * q).sum(1)])
    if kernel == O.KERNEL_REF_VEL3D:
        return np.array([(q * g / r2 * d[..., k]).sum(1) for k in range(3)])
    out = [(g / np.sqrt(r2) * q).sum(1)]
    out += [(-g * d[..., k] / r2 ** 1.5 * q).sum(1) for k in range(3)]
    return np.array(out)


@pytest.mark.parametrize("dim,kernel,conv,eps", [(3, O.KERNEL_POT, 0, 1e-3), (3, O.KERNEL_REF_VEL3D, 0, 1e-2),
                                                 (3, O.KERNEL_POT_GRAD, 1, 0.0), (2, O.KERNEL_POT, 0, 1e-3),
                                                 (2, O.KERNEL_POT_GRAD, 1, 0.0)])
def test_single_pair_interactions_match_closed_form(dim, kernel, conv, eps):
    x, y, z, q = wl.uniform_cube(3000, dim, 9)
    thresh, deg = 50, 4
    t = O.OracleTree(dim, thresh, x, y, z)
    cfg = O.make_config(dim, deg, thresh, 0.7, eps)
    o = O.opts(kernel=kernel, convention=conv)
    W = O.upward_pass(cfg, t, q)
    lst = O.traversal(cfg, t, t)
    lo, hi, child, leaf = panels(t)
    co = np.stack([x, y] + ([z] if dim == 3 else []), 1)
    P = (deg + 1) ** dim
    Cn = O.components(dim, kernel)
    for ty in range(4):
        cand = lst[(lst["interact_type"] == ty) & (lst["target_panel"] != lst["source_panel"])]
        if len(cand) == 0:
            continue
        e = cand[len(cand) // 2: len(cand) // 2 + 1]
        T, S = int(e["target_panel"][0]), int(e["source_panel"][0])
        out, Psi = O.fmm_interactions(cfg, o, t, t, q, W, e)
        tpts = t.leaf_points[t.leaf_start[T]:t.leaf_start[T + 1]]
        spts = t.leaf_points[t.leaf_start[S]:t.leaf_start[S + 1]]
        tg = np.stack(grid(lo[T], hi[T], deg, dim), 1)
        sg = np.stack(grid(lo[S], hi[S], deg, dim), 1)
        if ty == 0:
            ref, got = G(dim, kernel, conv, eps, co[tpts], co[spts], q[spts]), out[:, tpts]
        elif ty == 1:
            ref, got = G(dim, kernel, conv, eps, co[tpts], sg, W[S]), out[:, tpts]
        elif ty == 2:
            ref, got = G(dim, kernel, conv, eps, tg, co[spts], q[spts]), Psi[:, T]
        else:
            ref, got = G(dim, kernel, conv, eps, tg, sg, W[S]), Psi[:, T]
        assert got.shape == (Cn, ref.shape[1])
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_kernel_signs_and_values():
    one = np.array([1.0])
    o = O.opts(kernel=O.KERNEL_POT)
    # 3D reference: -1/(4 pi r) (poisson_fmm_impl.hpp:241)
    r = O.direct_sum(3, 0.0, o, one * 2, one * 0, one * 0, one * 0, one * 0, one * 0, one)
    assert r[0, 0] == pytest.approx(-1 / (8 * PI), rel=1e-15)
    # 2D reference: (1/4pi) log(r^2) (poisson_fmm_impl.hpp:38)
    r = O.direct_sum(2, 0.0, o, one * 2, one * 0, None, one * 0, one * 0, None, one)
    assert r[0, 0] == pytest.approx(np.log(4) / (4 * PI), rel=1e-15)
    # physical convention negates both
    op = O.opts(kernel=O.KERNEL_POT, convention=1)
    r = O.direct_sum(3, 0.0, op, one * 2, one * 0, one * 0, one * 0, one * 0, one * 0, one)
    assert r[0, 0] == pytest.approx(1 / (8 * PI), rel=1e-15)
    # eps regularisation, self term included: -1/(4 pi eps) (F3)
    r = O.direct_sum(3, 1e-4, o, one * 0, one * 0, one * 0, one * 0, one * 0, one * 0, one)
    assert r[0, 0] == pytest.approx(-1 / (4 * PI * 1e-4), rel=1e-14)
    # excluded self term
    oe = O.opts(kernel=O.KERNEL_POT, exclude_self=1)
    r = O.direct_sum(3, 0.0, oe, one * 0, one * 0, one * 0, one * 0, one * 0, one * 0, one)
    assert r[0, 0] == 0.0


@pytest.mark.parametrize("dim", [2, 3])
def test_gradient_is_derivative_of_potential(dim):
    x, y, z, q = wl.uniform_cube(200, dim, 13)
    rng = np.random.default_rng(1)
    t = rng.uniform(-0.5, 0.5, (5, dim)) + 3.0  # targets away from the sources
    og = O.opts(kernel=O.KERNEL_POT_GRAD, convention=1)
    op = O.opts(kernel=O.KERNEL_POT, convention=1)
    zt = t[:, 2].copy() if dim == 3 else None
    g = O.direct_sum(dim, 0.0, og, t[:, 0].copy(), t[:, 1].copy(), zt, x, y, z, q)
    h = 1e-5
    for k in range(dim):
        tp, tm = t.copy(), t.copy()
        tp[:, k] += h
        tm[:, k] -= h
        fp = O.direct_sum(dim, 0.0, op, tp[:, 0].copy(), tp[:, 1].copy(), tp[:, 2].copy() if dim == 3 else None,
                          x, y, z, q)[0]
        fm = O.direct_sum(dim, 0.0, op, tm[:, 0].copy(), tm[:, 1].copy(), tm[:, 2].copy() if dim == 3 else None,
                          x, y, z, q)[0]
        assert np.allclose(g[1 + k], (fp - fm) / (2 * h), rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("dim,deg", [(3, 5), (2, 8)])
def test_downward_interpolates_polynomials_exactly(dim, deg):
    x, y, z, q = wl.sphere_plummer_3d(4000, 4) if dim == 3 else wl.gaussian_mixture_2d(4000, 4)
    thresh = 60
    t = O.OracleTree(dim, thresh, x, y, z)
    cfg = O.make_config(dim, deg, thresh, 0.7, 0.0)
    lo, hi, child, leaf = panels(t)
    rng = np.random.default_rng(2)
    e = rng.integers(0, deg + 1, (3, dim))
    cf = rng.standard_normal(3)

    def f(pts):
        return sum(c * np.prod([pts[d] ** ee[d] for d in range(dim)], axis=0) for c, ee in zip(cf, e))

    P = (deg + 1) ** dim
    Psi = np.zeros((1, t.panel_count, P))
    Psi[0, 0] = f(grid(lo[0], hi[0], deg, dim))
    out = O.downward_pass(cfg, O.opts(kernel=O.KERNEL_POT), t, Psi, np.zeros((1, len(x))))
    co = [x, y, z]
    ref = f([co[d] for d in range(dim)])
    assert np.allclose(out[0], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_error_falls_with_degree():
    x, y, z, q = wl.uniform_cube(4000, 3, 17)
    ds = O.direct_sum(3, 0.0, O.opts(exclude_self=1), x, y, z, x, y, z, q)
    errs = []
    for deg in (2, 4, 6):
        cfg = O.make_config(3, deg, 100, 0.7, 0.0)
        out, _, _ = O.solve(cfg, O.opts(exclude_self=1), x, y, z, x, y, z, q)
        errs.append(np.linalg.norm(out - ds) / np.linalg.norm(ds))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-5


def _degree_matrix(d, e):
    """A[i][k] = L_k(y_i): the degree-d Lagrange basis on the nodes
    cos(pi k / d) at the degree-e proxy nodes cos(pi / e * i) of [-1, 1]."""
    a, b = np.cos(PI * np.arange(d + 1) / d), np.cos(PI / e * np.arange(e + 1))
    A = np.ones((e + 1, d + 1))
    for k in range(d + 1):
        for j in range(d + 1):
            if j != k:
                A[:, k] *= (b - a[j]) / (a[k] - a[j])
    return A


@pytest.mark.parametrize("d", [4, 6])
def test_degree_change_identities(d):
    """The two identities the solve with per-pair degrees stands on, on the
    oracle's own passes: the degree-d upward pass equals the degree-n one
    restricted by A^T in every dimension, and the degree-d downward pass of
    proxy potentials equals the degree-n one of the potentials extended by A
    (degree-d polynomials are reproduced by degree-n interpolation)."""
    n, thresh = 8, 60
    x, y, z, _ = wl.plummer_3d(2000, 4, a=0.3, rmax=3.0)
    q = np.random.default_rng(2).uniform(-1, 1, 2000)
    t = O.OracleTree(3, thresh, x, y, z)
    cn, cd = O.make_config(3, n, thresh, 0.7, 0.0), O.make_config(3, d, thresh, 0.7, 0.0)
    A = _degree_matrix(d, n)
    Wn = O.upward_pass(cn, t, q).reshape(-1, n + 1, n + 1, n + 1)
    Wd = O.upward_pass(cd, t, q)
    got = np.einsum("ak,bl,cm,pabc->pklm", A, A, A, Wn).reshape(Wd.shape)
    assert np.abs(got - Wd).max() <= 1e-13 * np.abs(Wd).max()
    o = O.opts(kernel=O.KERNEL_POT)
    Psi_d = np.random.default_rng(3).normal(size=(1, t.panel_count, (d + 1) ** 3))
    ref = O.downward_pass(cd, o, t, Psi_d, np.zeros((1, 2000)))
    Psi_n = np.einsum("ak,bl,cm,pklm->pabc", A, A, A, Psi_d[0].reshape(-1, d + 1, d + 1, d + 1))
    got = O.downward_pass(cn, o, t, Psi_n.reshape(1, t.panel_count, -1), np.zeros((1, 2000)))
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # the proxy nodes of degree n in place of its basis nodes' own degree-d
    # counterparts break the restriction: the identity is not vacuous
    A_bad = _degree_matrix(d, n)[::-1]
    bad = np.einsum("ak,bl,cm,pabc->pklm", A_bad, A, A, Wn).reshape(Wd.shape)
    assert np.abs(bad - Wd).max() > 1e-6 * np.abs(Wd).max()


def test_sampled_solve_reports_its_fractions():
    """The CPU baseline's sampling: strides of 1 run everything (fractions 1);
    larger strides run about one block of 64 units in `stride` of every part
    and report the fraction of each part's work they did."""
    x, y, z, q = wl.uniform_cube(15_000, 3, 5)
    cfg = O.make_config(3, 2, 8, 0.7, 0.0)
    O.solve(cfg, O.opts(kernel=O.KERNEL_POT_GRAD), x, y, z, x, y, z, q)
    assert O.last_near()[1] == 1.0 and O.last_far()[1] == 1.0
    assert np.all(O.last_passes()[1] == 1.0)
    O.solve(cfg, O.opts(kernel=O.KERNEL_POT_GRAD, near_stride=4, far_stride=5, pass_stride=3), x, y, z, x, y, z, q)
    assert abs(O.last_near()[1] - 1 / 4) < 0.08
    assert abs(O.last_far()[1] - 1 / 5) < 0.08
    sec, frac = O.last_passes()
    assert np.all(np.abs(frac - 1 / 3) < 0.12), frac
    assert np.all(sec >= 0)
