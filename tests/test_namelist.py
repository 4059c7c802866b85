"""read_run_config (read_namelist.cpp:10-44): the library's host reader against
the oracle's and against values derived by hand from the reference's rules.
tests/golden/namelist.txt is the reference's own namelist.txt (data file)."""
import os

import pytest

import barytreek_amd as btk
from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "namelist.txt")

FIELDS = ("interp_degree", "interp_point_count", "fmm_theta", "fmm_cluster_thresh", "ker_eps")

CASES = [
    # (text, dim, expected field values or None for "error")
    ("interp_degree=8\ncluster_size=100\ntheta=0.7\nkernel_eps=1e-4\nend", 3,
     dict(interp_degree=8, interp_point_count=729, fmm_theta=0.7, fmm_cluster_thresh=100, ker_eps=1e-4)),
    ("interp_degree=14\ncluster_size=0\nend\n", 2,          # degree clamped to 10; size <= 0 -> 2^dim * P
     dict(interp_degree=10, interp_point_count=121, fmm_cluster_thresh=484, fmm_theta=0.7, ker_eps=-1.0)),
    ("interp_degree=6\ncluster_size=-3\n", 3,              # EOF ends the parse
     dict(interp_degree=6, interp_point_count=343, fmm_cluster_thresh=2744, fmm_theta=0.7, ker_eps=-1.0)),
    ("theta=0.5\ninterp_degree = 9\ncluster_size=7\n", 3,  # ' =' is an unknown key: parse stops there
     dict(fmm_theta=0.5, ker_eps=-1.0)),
    ("kernel_eps=0\ntheta=1e-1xyz\nend", 2,               # trailing text after a number is ignored
     dict(ker_eps=0.0, fmm_theta=0.1)),
    ("interp_degree=3\ntheta\nend", 3,                     # no '=': value text of the previous line
     dict(interp_degree=3, fmm_theta=3.0)),
    ("interp_degree=abc\n", 3, None),                       # stoi throws
]


@pytest.mark.parametrize("text,dim,expect", CASES)
def test_reader_matches_oracle_and_rules(tmp_path, text, dim, expect):
    p = tmp_path / "namelist.txt"
    p.write_text(text)
    rc, oc = O.read_run_config(str(p), dim)
    if expect is None:
        assert rc != 0
        with pytest.raises(btk.BtkError):
            btk.read_run_config(str(p), dim)
        return
    assert rc == 0
    bc = btk.read_run_config(str(p), dim)
    for f, v in expect.items():
        assert getattr(bc, f) == pytest.approx(v, rel=0, abs=0), f
        assert getattr(oc, f) == pytest.approx(v, rel=0, abs=0), f
    for f in FIELDS:
        if f in expect:
            assert getattr(bc, f) == getattr(oc, f)


def test_reference_namelist_file():
    c = btk.read_run_config(GOLD, 3)
    assert (c.interp_degree, c.interp_point_count, c.fmm_cluster_thresh) == (8, 729, 100)
    assert c.fmm_theta == 0.7 and c.ker_eps == 1e-4


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(btk.BtkError, match="namelist not found"):
        btk.read_run_config(str(tmp_path / "nope.txt"), 3)
