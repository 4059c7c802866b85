"""Synthetic inputs of the benchmark configurations (seeded numpy generators).

The reference draws its points from std::random_device (unreproducible,
example_poisson_3d.cpp:44-45); here every generator takes an explicit seed.
"""
from __future__ import annotations

import numpy as np


def uniform_cube(n: int, dim: int, seed: int, charges: str = "uniform"):
    """n points uniform in [-1,1]^dim; charges U[-1,1] (or N(0,1))."""
    rng = np.random.default_rng(seed)
    pts = [rng.uniform(-1.0, 1.0, n) for _ in range(dim)]
    q = rng.uniform(-1.0, 1.0, n) if charges == "uniform" else rng.standard_normal(n)
    if dim == 2:
        pts.append(None)
    return pts[0], pts[1], pts[2], q


def gaussian_mixture_2d(n: int, seed: int, sigma: float = 0.05, frac_background: float = 0.2):
    """80% from 4 equal-weight isotropic Gaussians (centres U[-1,1]^2, sigma 0.05),
    20% uniform in [-1,1]^2; charges N(0,1)."""
    rng = np.random.default_rng(seed)
    nb = int(round(frac_background * n))
    ng = n - nb
    centres = rng.uniform(-1.0, 1.0, (4, 2))
    which = rng.integers(0, 4, ng)
    g = centres[which] + sigma * rng.standard_normal((ng, 2))
    b = rng.uniform(-1.0, 1.0, (nb, 2))
    p = np.concatenate([g, b])
    p = p[rng.permutation(n)]
    q = rng.standard_normal(n)
    return np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1]), None, q


def plummer(n: int, a: float, rmax: float, rng) -> np.ndarray:
    """Plummer sphere of scale a truncated at r <= rmax: r = a / sqrt(u^(-2/3) - 1)."""
    out = np.empty((0, 3))
    while len(out) < n:
        m = int((n - len(out)) * 1.3) + 16
        u = rng.uniform(0.0, 1.0, m)
        u = u[u > 0]
        r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = r[r <= rmax]
        v = rng.standard_normal((len(r), 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        out = np.concatenate([out, v * r[:, None]])
    return out[:n]


def sphere_plummer_3d(n: int, seed: int, a: float = 0.3, rmax: float = 2.0):
    """Half on the unit sphere (phi U[0,2pi), theta U[0,pi], taken literally),
    half Plummer a=0.3 truncated at r=2; charges U[-1,1]."""
    rng = np.random.default_rng(seed)
    ns = n // 2
    phi = rng.uniform(0.0, 2.0 * np.pi, ns)
    th = rng.uniform(0.0, np.pi, ns)
    s = np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], axis=1)
    p = np.concatenate([s, plummer(n - ns, a, rmax, rng)])
    p = p[rng.permutation(n)]
    q = rng.uniform(-1.0, 1.0, n)
    return np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1]), np.ascontiguousarray(p[:, 2]), q


def plummer_3d(n: int, seed: int, a: float = 0.3, rmax: float = 10.0):
    """Plummer sphere, charges U[-1,1]."""
    rng = np.random.default_rng(seed)
    p = plummer(n, a, rmax, rng)
    q = rng.uniform(-1.0, 1.0, n)
    return np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1]), np.ascontiguousarray(p[:, 2]), q
