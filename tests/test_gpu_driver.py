"""The namelist-driven driver (barytreek_amd/btk_driver, over the C++ drop-in
layer) runs the reference's four examples end to end on the GPU."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "barytreek_amd", "btk_driver")
NAMELIST = os.path.join(ROOT, "tests", "golden", "namelist.txt")


@pytest.mark.parametrize("example,n,bound", [("poisson_3d", 20000, 1e-6), ("poisson_2d", 10000, 1e-6),
                                             ("poisson_vel_3d", 20000, 1e-4), ("poisson_2d_split", 10000, 1e-6)])
def test_driver_examples(example, n, bound):
    print(f"\n[driver] {example} n={n}", flush=True)
    r = subprocess.run([DRIVER, example, "--namelist", NAMELIST, "--n", str(n), "--seed", "5"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:], r.stderr[-400:], flush=True)
    assert r.returncode == 0
    counts = dict(re.findall(r"^(\w[\w ]*) count: (\d+)$", r.stdout, re.M))
    assert int(counts["interaction"]) == sum(int(counts[k]) for k in ("pp interaction", "pc interaction",
                                                                      "cp interaction", "cc interaction"))
    err = float(re.search(r"relative l2 error is (\S+)", r.stdout).group(1))
    assert err < bound


def test_driver_reads_and_writes_btkf(tmp_path):
    """--points reads a BTKF file; --save writes config, tree, list and the solution."""
    import numpy as np
    import torch

    import barytreek_amd as btk
    from barytreek_amd import fileio, workloads as wl
    print("\n[driver] btkf round trip", flush=True)
    x, y, z, q = wl.uniform_cube(8000, 3, 3)
    src, dst = tmp_path / "in.btk", tmp_path / "out.btk"
    fileio.write(str(src), {"x": x, "y": y, "z": z, "q": q})
    r = subprocess.run([DRIVER, "poisson_3d", "--namelist", NAMELIST, "--points", str(src), "--save", str(dst),
                        "--no-direct-sum"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-400:], flush=True)
    assert r.returncode == 0
    f = fileio.read(str(dst))
    assert np.array_equal(f["x"], x) and np.array_equal(f["q"], q)
    cfg = btk.read_run_config(NAMELIST, 3)
    rc = btk.RunConfig.from_buffer_copy(f["config"].tobytes())
    assert (rc.dim, rc.interp_degree, rc.fmm_cluster_thresh, rc.fmm_theta) == (3, 8, 100, 0.7)
    assert rc.fmm_interaction_count == len(f["pairs"]) // 3
    out = torch.zeros(8000, dtype=torch.float64, device="cuda")
    d = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    btk.solve(cfg, btk.options(), d[0], d[1], d[2], d[3], out)
    ref = out.cpu().numpy()
    assert np.linalg.norm(f["soln"] - ref) / np.linalg.norm(ref) < 1e-12
    t = btk.Tree(cfg, d[0], d[1], d[2])
    assert f["panels"].tobytes() == t.panels().tobytes()
    assert np.array_equal(f["point_leaf_panel"], t.point_leaf_panel())
    assert np.array_equal(f["pairs"].reshape(-1, 3), btk.InteractionList(cfg, t, t).pairs().view(np.int32).reshape(-1, 3))
