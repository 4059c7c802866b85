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
