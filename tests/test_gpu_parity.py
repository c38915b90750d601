"""GPU DP vs the CPU oracle: same join tree, same rows/costs bit for bit."""
import pytest

import gpuqo_amd

from helpers import dpsub_sparse_graphs, assert_same_plan, cfg, gpu_bits, oracle_bits, small_graphs
from gpuqo_amd import qgen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,g", small_graphs(), ids=lambda x: x if isinstance(x, str) else "")
@pytest.mark.parametrize("bicc", [0, 1])
def test_small_graphs(name, g, bicc):
    c = cfg("dpsub", bicc)
    want, _ = oracle_bits(g, c)
    got, _ = gpu_bits(g, c)
    assert_same_plan(got, want, f"{name} bicc={bicc}")


def test_star12_dpsub():
    g = qgen.query_star12()
    c = cfg("dpsub", 0)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, "star-12 dpsub")
    assert gst.pairs == ost.pairs == 11 * 2 ** 11


def test_snowflake20_mpdp():
    g = qgen.query_snowflake20()
    c = cfg("dpsub", 1)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, "snowflake-20 mpdp")
    assert gst.pairs == ost.pairs
    assert gst.sets == ost.sets


@pytest.mark.parametrize("bicc", [0, 1])
def test_star18_one_and_two_lanes_per_edge(bicc):
    """Star-18: levels 8-12 (12,376 to 24,310 sets) exceed 4096 waves at two
    lanes per edge and take the one-lane form, the other levels the two-lane
    form; DPsub (bicc 0) ranks splits by pext, MPDP (bicc 1) by post-order."""
    g = qgen.star(18, 0x1818)
    c = cfg("dpsub", bicc)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, f"star-18 bicc={bicc}")
    assert gst.pairs == ost.pairs == 17 * 2 ** 17
    assert gst.sets == ost.sets


def test_clique14_dpsub():
    g = qgen.clique(14, 0x77)
    c = cfg("dpsub", 0)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, "clique-14 dpsub")
    assert gst.pairs == 3 ** 14 - 2 ** 15 + 1


@pytest.mark.parametrize("seed", range(4))
def test_mixed_batch_members(seed):
    for g in qgen.mixed_batch(6, 1000 + 97 * seed, nmin=8, nmax=16):
        c = cfg("dpsub", 1)
        want, _ = oracle_bits(g, c)
        got, _ = gpu_bits(g, c)
        assert_same_plan(got, want, f"{g.shape}{g.n}")


@pytest.mark.parametrize("name,g", small_graphs(), ids=lambda x: x if isinstance(x, str) else "")
def test_dpsize_small_graphs(name, g):
    c = cfg("dpsize")
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, f"{name} dpsize")
    assert gst.pairs == ost.pairs


def test_dpsize_star12():
    g = qgen.query_star12()
    c = cfg("dpsize")
    want, _ = oracle_bits(g, c)
    got, _ = gpu_bits(g, c)
    assert_same_plan(got, want, "star-12 dpsize")


def test_dpsize_clique12():
    g = qgen.clique(12, 0x55)
    c = cfg("dpsize")
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, "clique-12 dpsize")
    assert gst.pairs == ost.pairs == 3 ** 12 - 2 ** 13 + 1


def test_dpsize_ties_follow_creation_order():
    """All relations and selectivities identical: ties everywhere, so the
    plan is fixed by cpu_dpsize's list order and pair order alone."""
    g = qgen.QueryGraph(7, "ties")
    g.rows = [1000.0] * 7
    for i in range(7):
        for j in range(i + 1, 7):
            g.add_generic(i, j, 0.5)
    g.build()
    for algo in ("dpsize", "dpsub"):
        c = cfg(algo)
        want, _ = oracle_bits(g, c)
        got, _ = gpu_bits(g, c)
        assert_same_plan(got, want, f"ties {algo}")


@pytest.mark.parametrize("n,seed", [(40, 1), (70, 2), (130, 3)])
def test_uniondp_mpdp_large_snowflakes(n, seed):
    g = qgen.snowflake(n, seed, depth=4, fanout=5)
    c = cfg("dpsub", 1, idp_type="union", idp_n_iters=25)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, f"uniondp snowflake-{n}")
    assert gst.n_dp_calls == ost.n_dp_calls


@pytest.mark.parametrize("k", [5, 8, 12])
def test_idp2_snowflake20(k):
    g = qgen.snowflake(20, 4)
    c = cfg("dpsub", 1, idp_type="idp2", idp_n_iters=k)
    want, _ = oracle_bits(g, c)
    got, _ = gpu_bits(g, c)
    assert_same_plan(got, want, f"idp2 k={k}")


@pytest.mark.parametrize("bicc", [1, 0])
def test_batch_matches_oracle_and_single(bicc):
    """A mixed batch in one fused launch: every plan equals the oracle's and
    the plan of the same query run alone."""
    import gpuqo_amd
    qs = qgen.mixed_batch(40, 4242 + bicc, nmin=4, nmax=15)
    c = cfg("dpsub", bicc)
    res = gpuqo_amd.plan_batch(qs, c)
    for g, (nodes, st) in zip(qs, res):
        got = gpuqo_amd.plan_bits(nodes)
        want, _ = oracle_bits(g, c)
        assert_same_plan(got, want, f"batch {g.shape}{g.n}")
        alone, _ = gpu_bits(g, c)
        assert_same_plan(got, alone, f"alone {g.shape}{g.n}")


def test_fused_and_per_level_paths_agree():
    """Queries just above and below the fused-kernel size limit."""
    import gpuqo_amd
    for n in (13, 14, 15, 16):
        for g in (qgen.snowflake(n, 900 + n), qgen.star(n, 800 + n), qgen.cycle(n, 700 + n)):
            c = cfg("dpsub", 1)
            want, _ = oracle_bits(g, c)
            got, _ = gpu_bits(g, c)
            assert_same_plan(got, want, f"{g.shape}{n}")


def _big_block_graphs():
    return [("cycle18", qgen.cycle(18, 0x1801)), ("cycle22", qgen.cycle(22, 0x2201)),
            ("cycle16+tail4", qgen.cycle_with_tail(16, 4, 0x1604)),
            ("chorded15", qgen.chorded_cycle(15, 5, 0x1505)), ("chorded17", qgen.chorded_cycle(17, 3, 0x1717))]


@pytest.mark.parametrize("name,g", _big_block_graphs(), ids=lambda x: x if isinstance(x, str) else "")
def test_mpdp_big_blocks(name, g):
    """MPDP sets whose biconnected block has 15+ relations take the
    chunk-parallel big-set path."""
    c = cfg("dpsub", 1)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, name)
    assert gst.pairs == ost.pairs
    assert gst.sets == ost.sets


def test_batch_with_big_blocks():
    """Big-block MPDP sets of several queries in one batch."""
    import gpuqo_amd
    qs = [qgen.cycle(16, 0x1610), qgen.star(12, 0x1211), qgen.cycle_with_tail(16, 3, 0x1612),
          qgen.cycle(20, 0x2013), qgen.snowflake(14, 0x1414), qgen.chain(17, 0x1715), qgen.cycle(15, 0x1516)]
    c = cfg("dpsub", 1)
    res = gpuqo_amd.plan_batch(qs, c)
    for g, (nodes, st) in zip(qs, res):
        want, ost = oracle_bits(g, c)
        assert_same_plan(gpuqo_amd.plan_bits(nodes), want, f"batch {g.shape}{g.n}")
        assert st.pairs == ost.pairs, f"{g.shape}{g.n}"


@pytest.mark.parametrize("name,g", dpsub_sparse_graphs(), ids=[c[0] for c in dpsub_sparse_graphs()])
def test_dpsub_sparse_cyclic_through_blocks(name, g):
    """Plain DPsub on sparse cyclic graphs of 12 or more relations finds the
    valid splits of a set through its biconnected blocks and ranks them by
    DPsub's j, instead of testing all 2^k - 2 subsets: plan, pairs and sets
    equal cpu_dpsub's.  The triangle ring is one 16-relation block and keeps
    the subset scan."""
    c = cfg("dpsub", 0)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, f"dpsub {name}")
    assert (gst.pairs, gst.sets) == (ost.pairs, ost.sets)


def test_dpsub_sparse_cyclic_batched():
    """The same through gq_plan_batch, beside DPsub jobs that keep the subset
    scan and tree jobs."""
    import gpuqo_amd
    qs = [g for _, g in dpsub_sparse_graphs()] + [qgen.clique(10, 0xb30), qgen.star(13, 0xb31), qgen.cycle(9, 0xb32)]
    c = cfg("dpsub", 0)
    for g, (nodes, st) in zip(qs, gpuqo_amd.plan_batch(qs, c)):
        want, ost = oracle_bits(g, c)
        assert_same_plan(gpuqo_amd.plan_bits(nodes), want, f"batch dpsub {g.shape}{g.n}")
        assert st.pairs == ost.pairs, f"{g.shape}{g.n}"


def _idp1_cases():
    return [("snow16-mpdp", qgen.snowflake(16, 0x16), "dpsub", 1, 5),
            ("snow16-dpsub", qgen.snowflake(16, 0x16), "dpsub", 0, 4),
            ("snow20-mpdp", qgen.query_snowflake20(), "dpsub", 1, 8),
            ("cycle12-mpdp", qgen.cycle(12, 0x12), "dpsub", 1, 4),
            ("chain15-mpdp", qgen.chain(15, 0x15), "dpsub", 1, 6),
            ("clique10-dpsub", qgen.clique(10, 0x10), "dpsub", 0, 4),
            ("star12-dpsize", qgen.star(12, 0x12), "dpsize", 0, 5)]


@pytest.mark.parametrize("name,g,algo,bicc,k", _idp1_cases(), ids=lambda x: x if isinstance(x, str) else "")
def test_idp1(name, g, algo, bicc, k):
    """IDP1: DPs capped at k levels, each returning its cheapest k-relation
    plan, which becomes one relation of the next DP."""
    c = cfg(algo, bicc, idp_type="idp1", idp_n_iters=k)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, f"idp1 {name} k={k}")
    assert gst.pairs == ost.pairs


def _uniform(n, edges, shape):
    """Identical relations and selectivities: many sets of one size cost the
    same, so IDP1 must order ties as the CPU DP's memo does."""
    g = qgen.QueryGraph(n, shape)
    for i in range(n):
        g.rows[i] = qgen.rows_of(1e3)
    for (u, v) in edges:
        g.add_generic(u, v, 1e-3)
    return g.build()


@pytest.mark.parametrize("algo,bicc", [("dpsub", 1), ("dpsub", 0), ("dpsize", 0)])
def test_idp1_ties(algo, bicc):
    star = _uniform(9, [(0, i) for i in range(1, 9)], "uniform-star")     # C(8,2) = 28 tied 3-sets
    chain = _uniform(10, [(i, i + 1) for i in range(9)], "uniform-chain")  # 7 tied 4-paths
    big = _uniform(12, [(0, i) for i in range(1, 12)], "uniform-star")    # C(11,4) = 330 tied 5-sets
    for name, g, k in (("star9", star, 3), ("chain10", chain, 4), ("star12", big, 5)):
        c = cfg(algo, bicc, idp_type="idp1", idp_n_iters=k)
        want, _ = oracle_bits(g, c)
        got, _ = gpu_bits(g, c)
        assert_same_plan(got, want, f"idp1 ties {name} {algo} bicc={bicc}")


@pytest.mark.parametrize("kind", [2, 3], ids=["hash", "rank"])
def test_idp1_tree_memos(kind):
    """IDP1 on a tree DP with the open-addressing (2) or rank (3) memo."""
    g = qgen.snowflake(22, 0x22)
    c = cfg("dpsub", 1, idp_type="idp1", idp_n_iters=8, max_memo_size_mb=16, memo_kind=kind)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, f"idp1 snowflake-22 memo {kind}")
    assert gst.pairs == ost.pairs


SPANNING = [("cycle20", qgen.cycle(20, 0x5a01)), ("chorded22", qgen.chorded_cycle(22, 5, 0x5a02)),
            ("clique12", qgen.clique(12, 0x5a03)), ("cycle_tail", qgen.cycle_with_tail(9, 8, 0x5a04))]


@pytest.mark.parametrize("name,g", SPANNING, ids=[s[0] for s in SPANNING])
@pytest.mark.parametrize("algo,bicc", [("dpsub", 1), ("dpsub", 0), ("dpsize", 0)], ids=["mpdp", "dpsub", "dpsize"])
def test_spanning_tree_mode(name, g, algo, bicc):
    """gpuqo_spanning_tree: the DP runs on the minimum spanning tree, a tree
    graph whose classes are not all tree edges (the tree kernels take the
    general crossing-class test)."""
    if algo == "dpsize" and g.n > 14:
        pytest.skip("DPsize oracle at this size takes minutes")
    c = cfg(algo, bicc, spanning_tree=1)
    want, _ = oracle_bits(g, c)
    got, st = gpu_bits(g, c)
    assert_same_plan(got, want, f"{name} {algo} spanning tree")


@pytest.mark.parametrize("algo,bicc", [("dpsub", 1), ("dpsub", 0), ("dpsize", 0)], ids=["mpdp", "dpsub", "dpsize"])
def test_spanning_tree_mode_nan_selectivity_edge(algo, bicc):
    """An edge whose selectivity is NaN weighs 1 row in Prim's tree (the row
    estimate clamps NaN to 1) and is taken; the GPU path builds the same
    tree and plan as the oracle."""
    from test_oracle import nan_triangle
    g = nan_triangle()
    c = cfg(algo, bicc, spanning_tree=1)
    want, _ = oracle_bits(g, c)
    got, _ = gpu_bits(g, c)
    assert_same_plan(got, want, f"nan triangle {algo} spanning tree")


@pytest.mark.parametrize("kind", [2, 3], ids=["hash", "rank"])
def test_spanning_tree_mode_tree_memos(kind):
    g = qgen.chorded_cycle(24, 5, 0x5a05)
    c = cfg("dpsub", 1, spanning_tree=1, memo_kind=kind)
    want, _ = oracle_bits(g, c)
    got, st = gpu_bits(g, c)
    assert st.memo_kind == kind
    assert_same_plan(got, want, f"chorded24 spanning tree memo {kind}")


@pytest.mark.parametrize("idp,k", [("idp2", 10), ("union", 25), ("idp1", 8)])
def test_spanning_tree_mode_under_idp(idp, k):
    g = qgen.chorded_cycle(40, 7, 0x5a06) if idp == "union" else qgen.chorded_cycle(22, 5, 0x5a07)
    c = cfg("dpsub", 1, spanning_tree=1, idp_type=idp, idp_n_iters=k)
    want, _ = oracle_bits(g, c)
    got, _ = gpu_bits(g, c)
    assert_same_plan(got, want, f"{idp} k={k} spanning tree")


def test_spanning_tree_mode_batched():
    qs = [g for _, g in SPANNING] + [qgen.cycle(14, 0x5a08 + i) for i in range(6)]
    c = cfg("dpsub", 1, spanning_tree=1)
    for g, (nodes, _) in zip(qs, gpuqo_amd.plan_batch(qs, c)):
        want, _ = oracle_bits(g, c)
        assert_same_plan(gpuqo_amd.plan_bits(nodes), want, f"{g.shape}{g.n} batched spanning tree")


def test_dpsub_tree_flag_is_accepted_and_changes_no_plan():
    """gpuqo_dpsub_tree only selects the reference's per-edge enumeration of
    tree sets; here tree sets are always costed per edge in the CPU DPsub
    split order, so the flag changes no plan."""
    for g in (qgen.snowflake(16, 0x5c01), qgen.cycle(12, 0x5c02), qgen.chorded_cycle(14, 4, 0x5c03)):
        for bicc in (0, 1):
            want, _ = oracle_bits(g, cfg("dpsub", bicc))
            got, _ = gpu_bits(g, cfg("dpsub", bicc, dpsub_tree=1))
            assert_same_plan(got, want, f"{g.shape}{g.n} bicc={bicc} dpsub_tree")


@pytest.mark.parametrize("idp,k", [("union", 25), ("idp2", 12)])
def test_spanning_tree_mode_past_64_relations(idp, k):
    """80 relations: the query's sets take more than one 64-bit word on the
    host, so the spanning tree is built over the wide set type."""
    g = qgen.chorded_cycle(80, 9, 0x5d50)
    c = cfg("dpsub", 1, spanning_tree=1, idp_type=idp, idp_n_iters=k)
    want, _ = oracle_bits(g, c)
    got, _ = gpu_bits(g, c)
    assert_same_plan(got, want, f"chorded80 {idp} k={k} spanning tree")


def _nan_base(g, rel):
    g.build()
    g.base[rel]["total"] = float("nan")
    return g


@pytest.mark.parametrize("name,g", [
    ("snowflake12", _nan_base(qgen.snowflake(12, 61), 5)),
    ("star10", _nan_base(qgen.star(10, 62), 0)),
    ("snowflake22", _nan_base(qgen.snowflake(22, 63), 9)),
    ("chain16", _nan_base(qgen.chain(16, 64), 7)),
])
def test_tree_nan_costs(name, g):
    """A base relation with a NaN total makes every plan containing it cost
    NaN, so the tree evaluators' first-valid rule (the lowest-ranked split
    wins when its cost is NaN) decides those sets, in the two-lane levels,
    the one-lane levels and the batch kernel."""
    c = cfg("dpsub", 1)
    want, _ = oracle_bits(g, c)
    got, _ = gpu_bits(g, c)
    assert_same_plan(got, want, f"{name} NaN base cost")
    batch = gpuqo_amd.plan_batch([g, qgen.snowflake(14, 65)], c)
    assert_same_plan(gpuqo_amd.plan_bits(batch[0][0]), want, f"{name} NaN base cost, batched")
