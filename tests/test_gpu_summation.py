"""How mixed precision keeps its PP/PC sums (btk_options.summation), checked
target by target against the oracle on dense Plummer-core leaves of up to
1300 points: FP32 partial sums over blocks of 64 sources added into FP64
(BTK_SUM_BLOCKED), and the same blocks summed with Kahan's compensation
(BTK_SUM_KAHAN).

Both variants evaluate the same list as the oracle's FP64 treecode, so their
difference from it is rounding only.  For the potential (a positive kernel
times the charges) the rounding of a sum is bounded by a multiple of
2^-24 times the target's absolute sum A_t = sum_s |q_s| / (4 pi r_ts), taken
from the FP64 direct sum with |q|; every target must meet
|phi_t - phi_t(oracle)| <= 1e-6 A_t, and indeed <= 2^-23 A_t, for both
variants."""
import numpy as np
import pytest

import barytreek_amd as btk
from barytreek_amd import workloads as wl
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

U = 2.0 ** -24


def test_blocked_and_kahan_sums_per_target():
    n, leaf, deg = 60000, 1300, 8
    print(f"\n[summation] dense Plummer core, n={n} leaf={leaf} degree={deg}", flush=True)
    x, y, z, q = wl.plummer_3d(n, 61, a=0.1, rmax=1.0)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    cfg = btk.make_config(3, deg, leaf, 0.7, 0.0)
    C = btk.components(3, btk.KERNEL_POT_GRAD)
    out = {}
    for sm in (btk.SUM_BLOCKED, btk.SUM_KAHAN):
        o = torch.zeros(C * n, dtype=torch.float64, device="cuda")
        st = btk.solve(cfg, btk.options(kernel=btk.KERNEL_POT_GRAD, exclude_self=1, precision=btk.PREC_MIXED,
                                        summation=sm), d[0], d[1], d[2], d[3], o)
        out[sm] = o.view(C, n).cpu().numpy()
        print(f"  summation {sm}: near {st.t_kernel_near * 1e3:.2f} ms; pairs {list(st.count)}", flush=True)
    absq = torch.from_numpy(np.abs(q)).cuda()
    A = torch.zeros(n, dtype=torch.float64, device="cuda")
    btk.direct_sum(3, btk.options(kernel=btk.KERNEL_POT, exclude_self=1), 0.0, d[0], d[1], d[2], d[0], d[1], d[2],
                   absq, A)
    A = np.abs(A.cpu().numpy())
    O.set_threads(16)
    ref, _, _ = O.solve(O.make_config(3, deg, leaf, 0.7, 0.0), O.opts(kernel=O.KERNEL_POT_GRAD, exclude_self=1),
                        x, y, z, x, y, z, q)
    worst = {}
    for sm, name in ((btk.SUM_BLOCKED, "blocked"), (btk.SUM_KAHAN, "kahan")):
        g = out[sm]
        e_pot = np.abs(g[0] - ref[0]) / A
        rel = [float(np.linalg.norm(g[c] - ref[c]) / np.linalg.norm(ref[c])) for c in range(C)]
        worst[name] = float(e_pot.max())
        print(f"  {name}: potential |err| / A_t: max {e_pot.max():.3g} ({e_pot.max() / U:.1f} u), 99.9% "
              f"{np.quantile(e_pot, 0.999):.3g}, median {np.median(e_pot):.3g}; rel-L2 per component "
              f"{', '.join(f'{r:.2e}' for r in rel)}", flush=True)
        assert e_pot.max() <= 1e-6, name
        assert e_pot.max() <= 2 * U, name  # measured: 0.09 u (blocked), 0.03 u (Kahan)
        assert max(rel) < 1e-6, name
    # compensation cannot make the worst target's rounding grow by more than noise
    assert worst["kahan"] <= 1.5 * worst["blocked"]
