"""Single DPs over 33-64 relations: the wide DP (64-bit device sets, level
lists built by expansion, open-addressing memo), against fixtures the
oracle wrote (tests/golden/plans_wide_v1.json, `make_golden.py --wide`).

The oracle's DPsub and MPDP pass over all 2^n integers, which does not end
above about 32 relations; the fixtures come from oracle.plan_connected,
which visits only the connected sets, size by size.  CPU: that enumeration
gives the same plans and counters as the 2^n pass on small graphs of every
shape (trees, cycles, cliques, chorded cycles, IDP2 and UnionDP runs); the
generators still produce the recorded inputs; the fixtures' pair and set
counts are the subtree polynomial's; the oracle reproduces the small
fixtures.  GPU: gq_plan reproduces every fixture bit for bit with equal
counters; 64 relations (one more than the oracle's sets hold) against the
chain's closed forms; the unsupported cases fail with their status."""
import json
import os

import pytest

import gpuqo_amd
import oracle
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan, tree_pairs_and_sets

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "plans_wide_v1.json")) as f:
    WIDE = {r["label"]: r for r in json.load(f)["records"]}
LABELS = sorted(WIDE)


def _digest(g):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.digest(g)


def wide_fixture(label):
    r = WIDE[label]
    g = getattr(qgen, r["generator"])(*r["args"], **r["kwargs"])
    want = [(tuple(x), a, b, c, w, l, rr) for (x, a, b, c, w, l, rr) in r["plan"]]
    return g, want, r


def full_dp(r):
    return "idp_type" not in r["config"]


# ------------------------------------------------------------------ CPU


def _small_cases():
    out = []
    for n, s in ((9, 1), (12, 2), (16, 3), (18, 4)):
        out.append((f"snowflake{n}", qgen.snowflake(n, s, depth=5, fanout=3)))
    out += [("star10", qgen.star(10, 5)), ("chain14", qgen.chain(14, 6)), ("cycle12", qgen.cycle(12, 7)),
            ("clique8", qgen.clique(8, 8)), ("cycle8+tail6", qgen.cycle_with_tail(8, 6, 9)),
            ("chorded12", qgen.chorded_cycle(12, 3, 10)), ("hubchains15x4", qgen.hub_chains(15, 4, 11))]
    return out


@pytest.mark.parametrize("name,g", _small_cases(), ids=[n for n, _ in _small_cases()])
@pytest.mark.parametrize("bicc", [0, 1])
def test_connected_enumeration_equals_full_pass(name, g, bicc):
    """oracle_plan_connected against oracle_plan: plan bits, pairs, sets and
    the candidate (check) count."""
    c = abi.default_config(algorithm="dpsub", dpsub_bicc=bicc)
    a, b = oracle.plan(g, c), oracle.plan_connected(g, c)
    assert a[0] == b[0] == 0
    assert_same_plan(gpuqo_amd.plan_bits(b[1]), gpuqo_amd.plan_bits(a[1]), name)
    assert (b[2].pairs, b[2].sets, b[2].candidates) == (a[2].pairs, a[2].sets, a[2].candidates)


@pytest.mark.parametrize("over", [{"idp_type": "idp2", "idp_n_iters": 6}, {"idp_type": "union", "idp_n_iters": 10},
                                  {"idp_type": "idp1", "idp_n_iters": 5}],
                         ids=["idp2-k6", "union-k10", "idp1-k5"])
def test_connected_enumeration_equals_full_pass_under_idp(over):
    g = qgen.snowflake(30, 0x30c)
    c = abi.default_config(algorithm="dpsub", dpsub_bicc=1, **over)
    a, b = oracle.plan(g, c), oracle.plan_connected(g, c)
    assert a[0] == b[0] == 0
    assert_same_plan(gpuqo_amd.plan_bits(b[1]), gpuqo_amd.plan_bits(a[1]), str(over))
    assert (b[2].pairs, b[2].sets) == (a[2].pairs, a[2].sets)


def test_tree_pairs_and_sets_matches_brute_force():
    """The subtree polynomial against counting connected subsets directly."""
    for g in (qgen.snowflake(12, 3, depth=5, fanout=3), qgen.hub_chains(13, 3, 4), qgen.chain(11, 5)):
        adj = g.adjacency()
        sets = pairs = 0
        for s in range(1, 1 << g.n):
            k = bin(s).count("1")
            if k < 2:
                continue
            seen, todo = 0, s & -s
            while todo:
                v = (todo & -todo).bit_length() - 1
                todo &= todo - 1
                if not (seen >> v) & 1:
                    seen |= 1 << v
                    todo |= adj[v] & s & ~seen
            if seen == s:
                sets += 1
                pairs += 2 * (k - 1)
        assert tree_pairs_and_sets(g) == (pairs, sets)


def test_wide_fixture_inventory_and_inputs():
    assert len(WIDE) == 11
    for label in LABELS:
        g, want, r = wide_fixture(label)
        assert g.n == r["n"] and _digest(g) == r["input_sha256"], label
        assert 33 <= g.n <= 63 and len(want) == 2 * g.n - 1
        if full_dp(r):
            assert (r["pairs"], r["sets"]) == tree_pairs_and_sets(g), label
    chain = WIDE["chain63-mpdp"]
    assert chain["sets"] == 63 * 62 // 2 and chain["pairs"] == (63 ** 3 - 63) // 3


@pytest.mark.parametrize("label", ["chain63-dpsub", "chain63-mpdp", "snowflake34d6f2-mpdp"])
def test_oracle_reproduces_wide_fixture(label):
    """The fixtures that take the oracle seconds (the others take it
    minutes)."""
    g, want, r = wide_fixture(label)
    st, nodes, stats = oracle.plan_connected(g, abi.default_config(**r["config"]))
    assert st == 0
    assert_same_plan(gpuqo_amd.plan_bits(nodes), want, label)
    assert (stats.pairs, stats.sets) == (r["pairs"], r["sets"])


# ------------------------------------------------------------------ GPU


@pytest.mark.gpu
@pytest.mark.parametrize("label", LABELS)
def test_gpu_wide_matches_oracle_fixture(label):
    g, want, r = wide_fixture(label)
    nodes, st = gpuqo_amd.plan(g, abi.default_config(**r["config"]))
    assert_same_plan(gpuqo_amd.plan_bits(nodes), want, label)
    assert (st.pairs, st.sets) == (r["pairs"], r["sets"])
    if full_dp(r):
        assert st.memo_kind == abi.MEMO["hash"] and st.n_dp_calls == 1


def _plan_shape_ok(nodes, n):
    """Pre-order nodes of a join tree over relations 0..n-1: leaves are the
    relations, an inner node is the disjoint union of its children."""
    sets = [rel[0] for (rel, *_rest) in nodes]
    leaves = []
    for i, (rel, rows, su, tot, w, l, r) in enumerate(nodes):
        if l < 0:
            assert r < 0 and bin(sets[i]).count("1") == 1
            leaves.append(sets[i])
        else:
            assert sets[l] & sets[r] == 0 and sets[l] | sets[r] == sets[i]
            assert w == nodes[l][4] + nodes[r][4]
    assert sets[0] == (1 << n) - 1 and sorted(leaves) == [1 << i for i in range(n)]


@pytest.mark.gpu
@pytest.mark.parametrize("bicc", [0, 1])
def test_gpu_wide_64_relations(bicc):
    """64 relations, one more than the oracle's sets hold: a PK-FK chain
    (finite costs) against the chain's closed forms, and a hub of four
    chains against the subtree polynomial."""
    c = abi.default_config(algorithm="dpsub", dpsub_bicc=bicc)
    g = qgen.snowflake(64, 0x64, depth=63, fanout=1)
    nodes, st = gpuqo_amd.plan(g, c)
    assert (st.sets, st.pairs) == (64 * 63 // 2, (64 ** 3 - 64) // 3)
    _plan_shape_ok(nodes, 64)
    assert nodes[0][3] == nodes[0][3] and nodes[0][3] < float("inf")
    h = qgen.hub_chains(64, 4, 3)
    nodes, st = gpuqo_amd.plan(h, c)
    assert (st.pairs, st.sets) == tree_pairs_and_sets(h)
    _plan_shape_ok(nodes, 64)


@pytest.mark.gpu
def test_gpu_wide_in_a_batch_matches_alone():
    """gq_plan_batch plans its wide members one at a time: same plans."""
    qs = [qgen.snowflake(34, 0x34, depth=6, fanout=2), qgen.query_star12(), qgen.chain(40, 0x40)]
    c = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
    batch = gpuqo_amd.plan_batch(qs, c)
    for q, (nodes, _st) in zip(qs, batch):
        alone, _ = gpuqo_amd.plan(q, c)
        assert_same_plan(gpuqo_amd.plan_bits(nodes), gpuqo_amd.plan_bits(alone), q.shape)


def chord_block_graph():
    """34 relations: a 25-cycle with the chord (0, 12) -- one biconnected
    block of 25 that is not a simple cycle -- and a 9-relation tail."""
    edges = [(i, (i + 1) % 25) for i in range(25)] + [(0, 12)]
    prev = 0
    for i in range(25, 34):
        edges.append((prev, i))
        prev = i
    return qgen.graph(34, edges, 0x3425)


@pytest.mark.gpu
def test_gpu_wide_unsupported_cases():
    tree = qgen.snowflake(40, 0x40, depth=6, fanout=2)
    with pytest.raises(gpuqo_amd.GpuqoError) as e:          # a block of 25 that is not a cycle
        gpuqo_amd.plan(chord_block_graph(), abi.default_config(algorithm="dpsub", dpsub_bicc=1))
    assert e.value.status == abi.GQ_ERR_UNSUPPORTED and "biconnected block" in str(e.value)
    with pytest.raises(gpuqo_amd.GpuqoError) as e:          # DPsize above 32 relations
        gpuqo_amd.plan(tree, abi.default_config(algorithm="dpsize"))
    assert e.value.status == abi.GQ_ERR_UNSUPPORTED
    with pytest.raises(gpuqo_amd.GpuqoError) as e:          # memo budget
        gpuqo_amd.plan(tree, abi.default_config(algorithm="dpsub", dpsub_bicc=1, max_memo_size_mb=16))
    assert e.value.status == abi.GQ_ERR_NOMEM and "gpuqo_max_memo_size_mb" in str(e.value)
    # the library stays usable
    nodes, _ = gpuqo_amd.plan(qgen.query_star12(), abi.default_config(algorithm="dpsub"))
    assert len(nodes) == 23
