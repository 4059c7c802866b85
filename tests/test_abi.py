"""The C-ABI library loads and exports every symbol include/btk.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

import barytreek_amd as btk
from barytreek_amd import PANEL2D_DTYPE, PANEL3D_DTYPE, RunConfig, Options, SolveStats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "btk.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b(btk_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_header_matches_binding_list():
    assert declared_functions() == sorted(btk.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(btk.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_record_layouts_match_reference():
    # structs.hpp:6-21 (RunConfig 64 B), :29-44 (TreePanel_2d 80 B), :46-67 (TreePanel_3d 112 B)
    assert ctypes.sizeof(RunConfig) == 64
    assert RunConfig.fmm_theta.offset == 16 and RunConfig.ker_eps.offset == 48
    assert PANEL2D_DTYPE.itemsize == 80 and PANEL3D_DTYPE.itemsize == 112
    assert ctypes.sizeof(Options) == 32
    assert ctypes.sizeof(SolveStats) == 6 * 8 + 4 * 8 + 4 * 8 + 2 * 8 + 4 * 4 + 8
    assert SolveStats.cc_shared.offset == 144
    assert ctypes.sizeof(btk.Tuning) == 36


def test_tuning_round_trip_and_rejects_bad_values():
    """Host-only: the launch-shape record is stored and validated without a GPU."""
    old = btk.get_tuning()
    try:
        btk.set_tuning(near_items=32768, far_tpt=1)
        t = btk.get_tuning()
        assert (t.near_items, t.far_items, t.near_tpt, t.far_tpt) == (32768, 0, 0, 1)
        with pytest.raises(btk.BtkError, match="targets per lane"):
            btk.set_tuning(near_tpt=3)
        with pytest.raises(btk.BtkError, match="work item"):
            btk.set_tuning(far_items=-1)
        btk.set_tuning(near_waves=4, far_waves=2)
        t = btk.get_tuning()
        assert (t.near_waves, t.far_waves) == (4, 2)
        with pytest.raises(btk.BtkError, match="wavefronts per workgroup"):
            btk.set_tuning(near_waves=3)
        for v in (32, 64, 128):
            btk.set_tuning(multi_vectors=v)
            assert btk.get_tuning().multi_vectors == v
        with pytest.raises(btk.BtkError, match="charge vectors per workgroup"):
            btk.set_tuning(multi_vectors=96)
    finally:
        btk.restore_tuning(old)


def test_mac_costs_are_checked():
    btk.mac_set_costs(1.0, 1.0, 1.0, 0.35)
    for bad in ((0.0, 1.0, 1.0, 1.0), (1.0, -1.0, 1.0, 1.0), (1.0, 1.0, float("inf"), 1.0),
                (1.0, 1.0, 1.0, float("nan"))):
        with pytest.raises(btk.BtkError, match="positive"):
            btk.mac_set_costs(*bad)
    btk.mac_set_costs()


def test_components():
    assert btk.components(3, btk.KERNEL_POT) == 1
    assert btk.components(3, btk.KERNEL_REF_VEL3D) == 3
    assert btk.components(3, btk.KERNEL_POT_GRAD) == 4
    assert btk.components(2, btk.KERNEL_POT_GRAD) == 3
