"""The full-size workloads against fixtures the oracle wrote
(tests/golden/plans_large_v1.json, made by `make_golden.py --large`, which
takes the oracle about half an hour):

  * cpu_dpsub and cpu_dpsize on the 18-relation clique (153 classes);
  * cpu_dpsub_bicc on every query of the 256-query mixed batch
    (n in 8..22: 40% snowflake, 30% star, 15% chain, 15% cycle).

CPU: the generators still produce the recorded inputs, the clique's pair
count is the closed form 3^18 - 2^19 + 1, and the oracle reproduces a
sample of the batch records.  GPU: gq_plan (clique) and gq_plan_batch (the
whole batch in one call) reproduce every plan bit for bit, with equal pair
and set counts."""
import json
import os

import pytest

import gpuqo_amd
import oracle
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan, tree_pairs_and_sets

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "plans_large_v1.json")) as f:
    LARGE = json.load(f)["records"]
CLIQUE = [r for r in LARGE if r["generator"] == "query_clique18"]
BATCH = sorted((r for r in LARGE if r["generator"] == "queries_batch256"), key=lambda r: r["index"])
with open(os.path.join(HERE, "golden", "plans_trees_v1.json")) as f:
    TREES = {r["label"]: r for r in json.load(f)["records"]}
with open(os.path.join(HERE, "golden", "plans_served_v1.json")) as f:
    SERVED = sorted(json.load(f)["records"], key=lambda r: r["index"])

_MIX = None


def served_mix():
    global _MIX
    if _MIX is None:
        _MIX = qgen.serving_mix(256, qgen.workload_seed("serve256"))
    return _MIX


def tree_fixture(label):
    """(graph, expected plan bits, record) of a tree fixture."""
    r = TREES[label]
    g = getattr(qgen, r["generator"])(*r["args"], **r["kwargs"])
    return g, _want(r), r


def _digest(g):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.digest(g)


def _want(rec):
    return [(tuple(r), a, b, c, w, l, rr) for (r, a, b, c, w, l, rr) in rec["plan"]]


_BATCH_QS = None


def batch_queries():
    global _BATCH_QS
    if _BATCH_QS is None:
        _BATCH_QS = qgen.queries_batch256()
    return _BATCH_QS


def test_large_fixture_inventory():
    assert [r["label"] for r in CLIQUE] == ["clique18-dpsub", "clique18-dpsize"]
    assert [r["index"] for r in BATCH] == list(range(256))
    for r in CLIQUE:
        assert r["pairs"] == 3 ** 18 - 2 ** 19 + 1          # every split of a clique set is valid
        assert r["sets"] == 2 ** 18 - 18 - 1                 # every set of >= 2 relations is connected
        assert len(r["plan"]) == 2 * 18 - 1


def test_large_fixture_inputs_unchanged():
    assert _digest(qgen.query_clique18()) == CLIQUE[0]["input_sha256"] == CLIQUE[1]["input_sha256"]
    for q, r in zip(batch_queries(), BATCH):
        assert q.n == r["n"] and _digest(q) == r["input_sha256"], r["label"]


def test_tree_fixture_inventory_and_inputs():
    assert sorted(TREES) == ["snowflake24-mpdp", "snowflake26-mpdp", "snowflake28d4f5-mpdp", "snowflake30-mpdp"]
    for label, r in TREES.items():
        g, want, _ = tree_fixture(label)
        assert g.n == r["n"] and _digest(g) == r["input_sha256"], label
        assert len(want) == 2 * g.n - 1
        assert (r["pairs"], r["sets"]) == tree_pairs_and_sets(g), label


def test_served_fixture_inventory_and_inputs():
    """The served mix (config 5): 175 direct MPDP requests and 81 UnionDP
    (k = 25) requests, every one recorded with the configuration it is
    served with."""
    assert [r["index"] for r in SERVED] == list(range(256))
    for (g, large), r in zip(served_mix(), SERVED):
        assert g.n == r["n"] and _digest(g) == r["input_sha256"], r["label"]
        assert (r["config"].get("idp_type") == "union") == large, r["label"]
        assert len(r["plan"]) == 2 * g.n - 1
    assert sum(1 for r in SERVED if r["config"].get("idp_type") == "union") == 81


def test_oracle_reproduces_batch_sample():
    """The 24 smallest batch queries (the whole batch takes the oracle
    minutes)."""
    small = sorted(BATCH, key=lambda r: (r["n"], r["index"]))[:24]
    for r in small:
        g = batch_queries()[r["index"]]
        st, nodes, stats = oracle.plan(g, abi.default_config(**r["config"]))
        assert st == 0
        assert_same_plan(gpuqo_amd.plan_bits(nodes), _want(r), r["label"])
        assert stats.pairs == r["pairs"]


@pytest.mark.gpu
@pytest.mark.parametrize("rec", CLIQUE, ids=[r["label"] for r in CLIQUE])
def test_gpu_clique18_matches_oracle_fixture(rec):
    g = qgen.query_clique18()
    nodes, stats = gpuqo_amd.plan(g, abi.default_config(**rec["config"]))
    assert_same_plan(gpuqo_amd.plan_bits(nodes), _want(rec), rec["label"])
    assert stats.pairs == rec["pairs"] == 386896202
    assert stats.sets == rec["sets"]


@pytest.mark.gpu
def test_gpu_batch256_matches_oracle_fixtures():
    qs = batch_queries()
    res = gpuqo_amd.plan_batch(qs, abi.default_config(algorithm="dpsub", dpsub_bicc=1))
    total = 0
    for (nodes, stats), r in zip(res, BATCH):
        assert_same_plan(gpuqo_amd.plan_bits(nodes), _want(r), r["label"])
        assert stats.pairs == r["pairs"], r["label"]
        assert stats.sets == r["sets"], r["label"]
        total += stats.pairs
    assert total == sum(r["pairs"] for r in BATCH) == 470613398


@pytest.mark.gpu
def test_gpu_batch256_one_at_a_time_matches_oracle_fixtures():
    c = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
    for q, r in zip(batch_queries(), BATCH):
        nodes, stats = gpuqo_amd.plan(q, c)
        assert_same_plan(gpuqo_amd.plan_bits(nodes), _want(r), r["label"])
        assert stats.pairs == r["pairs"], r["label"]


@pytest.mark.gpu
def test_gpu_served_mix_alone_matches_oracle_fixtures():
    """Every request of the served mix planned alone through gq_plan (the
    UnionDP ones of 60-200 relations included) equals the oracle's plan."""
    for (g, _), r in zip(served_mix(), SERVED):
        nodes, stats = gpuqo_amd.plan(g, abi.default_config(**r["config"]))
        assert_same_plan(gpuqo_amd.plan_bits(nodes), _want(r), r["label"])
        assert stats.pairs == r["pairs"], r["label"]


@pytest.mark.gpu
def test_gpu_served_mix_through_the_server_matches_oracle_fixtures():
    """The whole served mix submitted at once to the continuous-batching
    server (96 GB arena, 128 request threads, as bench.py serves it): every
    plan equals the oracle's."""
    direct = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
    with gpuqo_amd.Server(direct, 96 << 30, max_inflight=128) as srv:
        for i, ((g, _), r) in enumerate(zip(served_mix(), SERVED)):
            srv.submit(g, i, abi.default_config(**r["config"]))
        got = {}
        while len(got) < len(SERVED):
            res = srv.poll(wait_ms=120000)
            assert res is not None, f"server returned nothing after {len(got)} requests"
            tag, st, nodes, stats = res
            assert st == 0, f"request {tag}: status {st}"
            got[tag] = (nodes, stats)
    for i, r in enumerate(SERVED):
        assert_same_plan(gpuqo_amd.plan_bits(got[i][0]), _want(r), r["label"] + " served")
        assert got[i][1].pairs == r["pairs"], r["label"]
