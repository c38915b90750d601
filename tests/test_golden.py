"""Committed golden plans (tests/golden/plans_v1.json, written by
tests/golden/make_golden.py): the generator still produces the same inputs,
the oracle still produces the same plans, and (GPU) so does the HIP path."""
import json
import os

import pytest

import gpuqo_amd
import oracle
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "plans_v1.json")) as f:
    GOLDEN = json.load(f)["records"]


def _load(rec):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    g = getattr(qgen, rec["generator"])(*rec["args"], **rec["kwargs"])
    return g, mg.digest(g)


def _want(rec):
    return [(tuple(r), a, b, c, w, l, rr) for (r, a, b, c, w, l, rr) in rec["plan"]]


@pytest.mark.parametrize("rec", GOLDEN, ids=[r["label"] for r in GOLDEN])
def test_oracle_reproduces_golden(rec):
    g, d = _load(rec)
    assert d == rec["input_sha256"], "generator output changed"
    st, nodes, stats = oracle.plan(g, abi.default_config(**rec["config"]))
    assert st == 0
    assert_same_plan(gpuqo_amd.plan_bits(nodes), _want(rec), rec["label"])
    assert stats.pairs == rec["pairs"]


@pytest.mark.gpu
@pytest.mark.parametrize("rec", GOLDEN, ids=[r["label"] for r in GOLDEN])
def test_gpu_matches_golden(rec):
    g, d = _load(rec)
    assert d == rec["input_sha256"], "generator output changed"
    nodes, stats = gpuqo_amd.plan(g, abi.default_config(**rec["config"]))
    assert_same_plan(gpuqo_amd.plan_bits(nodes), _want(rec), rec["label"])
    if "idp_type" not in rec["config"]:
        assert stats.pairs == rec["pairs"]
