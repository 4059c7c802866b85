"""Learned MAC on the GPU: the forward and gradient kernels against the CPU
restatement (oracle/mlp.py); the sample generator against a CPU
recomputation of the same pair errors (oracle basis and proxy nodes); and
what a trained network gains over the fixed theta MAC."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl
from oracle import mlp
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(autouse=True)
def _restore_model():
    try:
        old = btk.mac_get_weights()
    except btk.BtkError:
        old = None
    yield
    if old is not None:
        btk.mac_set_weights(old)
    btk.mac_set_budget(1e-7)


def test_forward_and_gradient_match_cpu_oracle():
    print("\n[mac] forward/gradient vs oracle", flush=True)
    rng = np.random.default_rng(11)
    w = rng.normal(0, 0.2, mlp.N_PARAMS).astype(np.float32)
    n = 300
    X = rng.uniform(0, 1, (n, mlp.N_IN)).astype(np.float32)
    Y = rng.normal(-1, 0.5, (n, mlp.N_OUT)).astype(np.float32)
    btk.mac_set_weights(w)
    dX, dY = torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()
    out = torch.zeros(n * mlp.N_OUT, dtype=torch.float32, device="cuda")
    btk.mac_forward(dX, n, out)
    ref, _ = mlp.forward(w.astype(np.float64), X.astype(np.float64))
    got = out.view(n, mlp.N_OUT).cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
    g = torch.zeros(mlp.N_PARAMS, dtype=torch.float32, device="cuda")
    loss = btk.mac_gradients(dX, dY, n, g)
    rl, rg = mlp.loss_and_grad(w.astype(np.float64), X.astype(np.float64), Y.astype(np.float64))
    assert abs(loss - rl) <= 1e-5 * rl
    gg = g.cpu().numpy()
    assert np.linalg.norm(gg - rg) / np.linalg.norm(rg) < 1e-4
    # every layer's block agrees, not just the total
    o = 0
    for name, shape in mlp.SIZES:
        k = int(np.prod(shape))
        assert np.linalg.norm(gg[o:o + k] - rg[o:o + k]) <= 1e-4 * np.linalg.norm(rg) + 1e-9, name
        o += k


def _charge(seed, i):
    """splitmix64 of (seed << 32) ^ i -> U[-1, 1), as the sample kernel draws it."""
    m = (1 << 64) - 1
    z = ((seed << 32) ^ int(i)) & m
    z = (z + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return (z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0


def _pair_errors(P, pts, start, T, S, seed, deg):
    """PC, CP, CC relative errors of the pair at degree deg (potential 1/r and
    gradient kernel d/r^3, L2 over the four components), restated on the CPU."""
    def picks(p, m):
        c = P["point_count"][p]
        return [start[p] + (k * c) // m for k in range(m)]
    ti = picks(T, min(64, P["point_count"][T]))
    si = picks(S, min(1024, P["point_count"][S]))
    t, s = pts[ti], pts[si]
    q = np.array([_charge(seed, i) for i in si])
    box = lambda p: (np.array([P["min_x"][p], P["min_y"][p], P["min_z"][p]]),
                     np.array([P["max_x"][p], P["max_y"][p], P["max_z"][p]]))
    def g(a, b):  # [4, len(a), len(b)]: 1/r and d/r^3
        d = a[:, None, :] - b[None, :, :]
        inv = 1.0 / np.linalg.norm(d, axis=2)
        return np.concatenate([inv[None], np.moveaxis(d, 2, 0) * inv[None] ** 3])
    exact = g(t, s) @ q
    (tl, th), (sl, sh) = box(T), box(S)

    def basis(x, lo, hi):  # tensor-product basis at points x: [len(x), P]
        B = [np.array([po.interp_vals_bli(v, lo[d], hi[d], deg) for v in x[:, d]]) for d in range(3)]
        return np.einsum("pa,pb,pc->pabc", B[0], B[1], B[2]).reshape(len(x), -1)

    def proxies(lo, hi):
        ax = [po.bli_points_shift(lo[d], hi[d], deg) for d in range(3)]
        return np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)

    W = basis(s, sl, sh).T @ q
    xs, xt = proxies(sl, sh), proxies(tl, th)
    Lt = basis(t, tl, th)
    pc = g(t, xs) @ W
    cp = np.einsum("tp,cp->ct", Lt, g(xt, s) @ q)   # every component interpolated on its own
    cc = np.einsum("tp,cp->ct", Lt, g(xt, xs) @ W)
    return [np.linalg.norm(a - exact) / np.linalg.norm(exact) for a in (pc, cp, cc)]


def test_samples_match_cpu_recomputation():
    print("\n[mac] samples vs CPU", flush=True)
    x, y, z, _ = wl.plummer_3d(20000, 5, a=0.3, rmax=3.0)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z)]
    cfg = btk.make_config(3, 8, 300, 0.7, 0.0)
    ns, seed = 5, 17
    X = torch.zeros(ns * 10, dtype=torch.float32, device="cuda")
    Y = torch.zeros(ns * 12, dtype=torch.float32, device="cuda")
    pr = torch.zeros(ns * 2, dtype=torch.int32, device="cuda")
    m = btk.mac_sample(cfg, d[0], d[1], d[2], ns, seed, X, Y, pr)
    assert m == ns
    tree = btk.Tree(cfg, d[0], d[1], d[2])
    P = tree.panels()
    perm, start = tree.permutation()
    pts = np.stack([x, y, z], 1)[perm]  # tree order
    Yh = Y.view(ns, 12).cpu().numpy() * 4.0
    Xh = X.view(ns, 10).cpu().numpy()
    for k, (T, S) in enumerate(pr.view(ns, 2).cpu().numpy()):
        ct, cs = P["point_count"][T], P["point_count"][S]
        cen = lambda p: 0.5 * np.array([P["min_x"][p] + P["max_x"][p], P["min_y"][p] + P["max_y"][p],
                                        P["min_z"][p] + P["max_z"][p]])
        dist = np.linalg.norm(cen(S) - cen(T))
        sep = (P["radius"][T] + P["radius"][S]) / dist
        assert abs(Xh[k][0] - min(sep, 4.0)) < 1e-5 and 0.3 <= sep < 1.2 and Xh[k][8] > 0
        assert abs(Xh[k][3] - np.log2(ct + 1) / 16) < 1e-6 and abs(Xh[k][4] - np.log2(cs + 1) / 16) < 1e-6
        for di, deg in enumerate((4, 6, 8, 10)):
            errs = _pair_errors(P, pts, start, T, S, seed, deg)
            for ty in range(3):
                ref = min(0.0, np.log10(max(errs[ty], 1e-14)))
                if ref > -12:
                    assert abs(Yh[k][ty * 4 + di] - ref) < 0.02, (k, deg, ty, Yh[k][ty * 4 + di], ref)


def test_trained_network_gains_over_the_theta_mac():
    """Train on samples of two Plummer sets, then solve a third: the learned
    MAC approximates more pairs (fewer kernel evaluations) and the result
    stays within 1e-6 of the FP64 direct sum."""
    print("\n[mac] training and gain", flush=True)
    cfg = btk.make_config(3, 8, 400, 0.7, 0.0)
    Xs, Ys, total = [], [], 0
    for seed in (1, 2):
        x, y, z, _ = wl.plummer_3d(150000, seed, a=0.3, rmax=10.0)
        d = [torch.from_numpy(a).cuda() for a in (x, y, z)]
        X = torch.zeros(6000 * 10, dtype=torch.float32, device="cuda")
        Y = torch.zeros(6000 * 12, dtype=torch.float32, device="cuda")
        m = btk.mac_sample(cfg, d[0], d[1], d[2], 6000, 100 + seed, X, Y)
        Xs.append(X[:m * 10])
        Ys.append(Y[:m * 12])
        total += m
    X, Y = torch.cat(Xs), torch.cat(Ys)
    loss = btk.mac_train(X, Y, total, epochs=150, batch=512, lr=2e-3, seed=7, reinit=True)
    print(f"  {total} samples, final loss {loss:.4f} (log10/4 units)", flush=True)
    # the samples' own scatter (where the points lie inside the boxes) leaves
    # about 0.1, a 1.3-decade rms error of the predicted log10
    assert loss < 0.15
    btk.mac_set_budget(2e-7)
    x, y, z, q = wl.plummer_3d(150000, 9, a=0.3, rmax=10.0)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    idx = torch.from_numpy(np.arange(0, 150000, 15).astype(np.int32)).cuda()
    ds = torch.zeros(idx.numel(), dtype=torch.float64, device="cuda")
    btk.direct_sum(3, btk.options(exclude_self=1), 0.0, d[0], d[1], d[2], d[0], d[1], d[2], d[3], ds, idx=idx)
    res = {}
    for mac in (0, 1):
        out = torch.zeros(150000, dtype=torch.float64, device="cuda")
        st = btk.solve(cfg, btk.options(exclude_self=1, mac=mac), d[0], d[1], d[2], d[3], out)
        err = float(torch.linalg.norm(out[idx.long()] - ds) / torch.linalg.norm(ds))
        res[mac] = (sum(st.evals[k] for k in range(4)), err, [int(st.count[k]) for k in range(4)])
        print(f"  mac={mac}: evals {res[mac][0]:.3e} rel {err:.2e} pairs {res[mac][2]}", flush=True)
    assert res[1][0] < 0.95 * res[0][0]  # fewer evaluations
    assert res[1][1] <= 1e-6             # within the metric's error
