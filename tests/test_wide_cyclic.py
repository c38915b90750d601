"""Single DPs over 33-64 relations whose join graphs have cycles: the wide
DP's general path (biconnected blocks of each set found on the device,
sets generated from every set they extend and kept once).  Fixtures from
the oracle's connected-sets enumeration (tests/golden/plans_wide_cyclic_v1.json,
`make_golden.py --wide-cyclic`): a 20-cycle with a 14-relation tail,
4-cycles joined by paths (36 relations), a hub of three chains with
triangles (40 relations), under DPsub and MPDP.

CPU: the generators still produce the recorded inputs and the oracle
reproduces every record (the oracle's connected enumeration is pinned
against its 2^n pass on small cyclic graphs in test_wide.py).  GPU:
gq_plan reproduces every record bit for bit with equal pair and set
counts; a 40-cycle (whose single block the oracle's subset enumeration
cannot finish) against the cycle's closed forms, with DPsub and MPDP at
the same optimum; a memo budget too small fails with GQ_ERR_NOMEM."""
import json
import os

import pytest

import gpuqo_amd
import oracle
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "plans_wide_cyclic_v1.json")) as f:
    CYC = {r["label"]: r for r in json.load(f)["records"]}
LABELS = sorted(CYC)


def _digest(g):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.digest(g)


def fixture(label):
    r = CYC[label]
    g = getattr(qgen, r["generator"])(*r["args"], **r["kwargs"])
    want = [(tuple(x), a, b, c, w, l, rr) for (x, a, b, c, w, l, rr) in r["plan"]]
    return g, want, r


def test_cyclic_fixture_inventory_and_inputs():
    assert len(CYC) == 6
    for label in LABELS:
        g, want, r = fixture(label)
        assert g.n == r["n"] and _digest(g) == r["input_sha256"], label
        assert 33 <= g.n <= 63 and len(want) == 2 * g.n - 1
        edges = sum(bin(a).count("1") for a in g.adjacency()) // 2
        assert edges > g.n - 1, label                        # not a tree
    # the 20-cycle with its tail: 3,160 connected sets (1 + 19 + ... arcs
    # through relation 0 times the tail's prefixes, plus the rest)
    assert CYC["cycle20tail14-mpdp"]["sets"] == 3160 - 34


@pytest.mark.parametrize("label", LABELS)
def test_oracle_reproduces_cyclic_fixture(label):
    g, want, r = fixture(label)
    st, nodes, stats = oracle.plan_connected(g, abi.default_config(**r["config"]))
    assert st == 0
    assert_same_plan(gpuqo_amd.plan_bits(nodes), want, label)
    assert (stats.pairs, stats.sets) == (r["pairs"], r["sets"])


@pytest.mark.gpu
@pytest.mark.parametrize("label", LABELS)
def test_gpu_cyclic_matches_oracle_fixture(label):
    g, want, r = fixture(label)
    nodes, st = gpuqo_amd.plan(g, abi.default_config(**r["config"]))
    assert_same_plan(gpuqo_amd.plan_bits(nodes), want, label)
    assert (st.pairs, st.sets) == (r["pairs"], r["sets"])
    assert st.memo_kind == abi.MEMO["hash"] and st.n_dp_calls == 1


def finite_cycle(n, seed):
    """A cycle of n relations of 20-200 rows with selectivities of 1e-3 to
    1e-2, so that every join's rows and costs stay finite."""
    import random
    r = random.Random(seed)
    g = qgen.QueryGraph(n, "cycle")
    for i in range(n):
        g.rows[i] = float(r.randint(20, 200))
    for i in range(n):
        g.add_generic(i, (i + 1) % n, qgen.f32(r.uniform(1e-3, 1e-2)))
    return g.build()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [40, 63, 64])
def test_gpu_cycle_closed_forms(n):
    """A cycle of n relations: n(n - 2) + 1 connected sets of two or more
    relations and n(n - 1)^2 csg-cmp pairs; DPsub and MPDP reach the same
    (finite) cost.  At 64 relations the top set's block (the whole cycle)
    leaves MPDP's ranks no room (2^64 - 2 subsets): DPsub only."""
    g = finite_cycle(n, 0x4040 + n)
    got = {}
    if n == 64:
        with pytest.raises(gpuqo_amd.GpuqoError) as e:
            gpuqo_amd.plan(g, abi.default_config(algorithm="dpsub", dpsub_bicc=1))
        assert e.value.status == abi.GQ_ERR_UNSUPPORTED
    for bicc in ((0,) if n == 64 else (0, 1)):
        nodes, st = gpuqo_amd.plan(g, abi.default_config(algorithm="dpsub", dpsub_bicc=bicc))
        assert (st.sets, st.pairs) == (n * (n - 2) + 1, n * (n - 1) ** 2)
        sets = [rel[0] for (rel, *_r) in nodes]
        assert sets[0] == (1 << n) - 1 and len(nodes) == 2 * n - 1
        assert 0 < nodes[0][3] < float("inf")
        got[bicc] = gpuqo_amd.fbits(nodes[0][3])
    assert len(set(got.values())) == 1


@pytest.mark.gpu
def test_gpu_cyclic_memo_budget():
    g, _want, r = fixture("squares36-mpdp")
    with pytest.raises(gpuqo_amd.GpuqoError) as e:
        gpuqo_amd.plan(g, abi.default_config(max_memo_size_mb=1, **r["config"]))
    assert e.value.status == abi.GQ_ERR_NOMEM and "gpuqo_max_memo_size_mb" in str(e.value)
    nodes, _ = gpuqo_amd.plan(g, abi.default_config(**r["config"]))      # still usable
    assert len(nodes) == 2 * g.n - 1
