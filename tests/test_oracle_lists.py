"""Pins the oracle's dual traversal (interaction_list.cpp:136-240) to
interaction lists derived by hand (tests/golden/tiny_lists.json): the exact
entry order, the types and the counters, for coincident and separate target
and source trees in 2D and 3D.

The cases include non-leaf pairs with equal point counts (the source must
split, interaction_list.cpp:196) and accepted pairs of every type.  Two
builds of the oracle with a planted mistake -- '>=' in the refine rule, and
the last two children pushed in swapped order -- must each fail at least one
case, which shows the fixtures can tell these rules apart.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as O

HERE = os.path.dirname(__file__)
LISTS = json.load(open(os.path.join(HERE, "golden", "tiny_lists.json")))
TREES = {t["name"]: t for t in json.load(open(os.path.join(HERE, "golden", "tiny_trees.json")))["trees"]}
TREES.update(LISTS["sources"])


def oracle_tree(name):
    t = TREES[name]
    z = np.array(t["z"]) if t["dim"] == 3 else None
    return O.OracleTree(t["dim"], t["thresh"], np.array(t["x"]), np.array(t["y"]), z)


def run_case(case):
    tt = oracle_tree(case["tree"])
    ts = oracle_tree(case["source"]) if "source" in case else tt
    dim, thresh = TREES[case["tree"]]["dim"], TREES[case["tree"]]["thresh"]
    cfg = O.make_config(dim, 2, thresh, case["theta"], 0.0)
    lst = O.traversal(cfg, tt, ts)
    got = [[int(e["target_panel"]), int(e["source_panel"]), int(e["interact_type"])] for e in lst]
    return got, cfg


def test_source_tree_is_as_derived():
    t = LISTS["sources"]["2d_five_sources"]
    ot = oracle_tree("2d_five_sources")
    assert ot.panel_count == t["panel_count"]
    assert list(ot.point_leaf_panel) == t["point_leaf_panel"]
    assert [ot.t.panels[i].point_count for i in range(ot.panel_count)] == t["point_count"]


@pytest.mark.parametrize("case", LISTS["cases"], ids=lambda c: c["name"])
def test_hand_derived_lists(case):
    got, cfg = run_case(case)
    assert got == case["list"]
    types = [e[2] for e in case["list"]]
    assert cfg.fmm_interaction_count == len(types)
    assert [cfg.fmm_pp_count, cfg.fmm_pc_count, cfg.fmm_cp_count, cfg.fmm_cc_count] == \
        [types.count(k) for k in range(4)]


def test_fixtures_cover_every_type_and_an_equal_count_split():
    seen = set()
    for case in LISTS["cases"]:
        seen |= {e[2] for e in case["list"]}
    assert seen == {0, 1, 2, 3}


@pytest.mark.parametrize("plant", ["ORC_PLANT_REFINE_GE", "ORC_PLANT_PUSH_ORDER"])
def test_planted_mistakes_fail(plant, tmp_path):
    src = os.path.join(os.path.dirname(HERE), "oracle", "btk_oracle.c")
    so = str(tmp_path / f"liboracle_{plant}.so")
    subprocess.run(["gcc", "-std=gnu11", "-O1", "-ffp-contract=off", "-fPIC", "-fopenmp", "-shared",
                    f"-D{plant}", "-o", so, src, "-lm"], check=True)
    prev = O.use_library(so)
    try:
        failed = [c["name"] for c in LISTS["cases"] if run_case(c)[0] != c["list"]]
    finally:
        O._lib = prev
    assert failed, f"{plant}: every hand-derived list still matches"
    # both planted mistakes change the 2D and the 3D cases
    assert any(n.startswith("2d") for n in failed) and any(n.startswith("3d") for n in failed), failed
