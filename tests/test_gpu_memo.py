"""The open-addressing and rank memos (tree graphs): the same plans as the
dense memo and the oracle, with level lists generated by expansion.  The rank
memo gives every connected set its own slot, computed from the set; slots
that collided or were skipped would break the plan or the set counts."""
import pytest

from helpers import assert_same_plan, cfg, gpu_bits, oracle_bits
from gpuqo_amd import qgen

pytestmark = pytest.mark.gpu

GQ_MEMO_DENSE, GQ_MEMO_HASH, GQ_MEMO_RANK = 1, 2, 3


def _trees():
    return [("snowflake20", qgen.query_snowflake20(), 1), ("star12", qgen.query_star12(), 0), ("chain15", qgen.chain(15, 7), 1),
            ("snow16-dpsub", qgen.snowflake(16, 33), 0), ("star9", qgen.star(9, 5), 1),
            ("chain2", qgen.chain(2, 1), 1), ("snow3", qgen.snowflake(3, 2), 0),
            ("snow18-d4", qgen.snowflake(18, 44, depth=4, fanout=5), 1)]


@pytest.mark.parametrize("kind", [GQ_MEMO_HASH, GQ_MEMO_RANK], ids=["hash", "rank"])
@pytest.mark.parametrize("name,g,bicc", _trees(), ids=lambda x: x if isinstance(x, str) else "")
def test_tree_memo_matches_oracle(name, g, bicc, kind):
    c = cfg("dpsub", bicc, memo_kind=kind)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, name)
    assert gst.memo_kind == kind
    assert gst.pairs == ost.pairs
    assert gst.sets == ost.sets


def test_rank_memo_on_stars_and_chains():
    """Extreme shapes of the slot numbering: a star (2^(n-1) sets with the
    centre as lowest member), a chain (n(n+1)/2 sets), and a star whose
    centre is not relation 0."""
    g3 = qgen.QueryGraph(9, "star-off-centre")
    for i in range(9):
        g3.rows[i] = qgen.rows_of(1e3 * (i + 1))
    for i in range(9):
        if i != 4:
            g3.add_generic(4, i, 1e-3 * (i + 1))
    g3.build()
    for g in (qgen.star(16, 0x516), qgen.chain(24, 0x524), g3):
        c = cfg("dpsub", 1, memo_kind=GQ_MEMO_RANK)
        want, ost = oracle_bits(g, c)
        got, gst = gpu_bits(g, c)
        assert_same_plan(got, want, g.shape)
        assert gst.sets == ost.sets


def test_auto_memo_switches_above_the_memo_limit():
    g = qgen.snowflake(22, 0x22)
    want, _ = oracle_bits(g, cfg("dpsub", 1))
    got, gst = gpu_bits(g, cfg("dpsub", 1, max_memo_size_mb=16))      # dense: 36 B x 2^22 = 151 MB
    assert gst.memo_kind == GQ_MEMO_RANK
    assert_same_plan(got, want, "auto -> rank")
    got2, gst2 = gpu_bits(g, cfg("dpsub", 1))                           # default limit 1024 MB
    assert gst2.memo_kind == GQ_MEMO_DENSE
    assert_same_plan(got2, want, "auto -> dense")


def test_hash_request_on_a_cycle_keeps_the_dense_memo():
    g = qgen.cycle(10, 3)
    c = cfg("dpsub", 1, memo_kind=GQ_MEMO_HASH)
    want, _ = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert gst.memo_kind == GQ_MEMO_DENSE
    assert_same_plan(got, want, "cycle10")


def test_hash_and_rank_equal_dense_at_30_relations():
    """All three memos against the oracle's plan of the 30-relation
    snowflake (tests/golden/plans_trees_v1.json)."""
    from test_golden_large import tree_fixture
    g, want, rec = tree_fixture("snowflake30-mpdp")
    dense, dst = gpu_bits(g, cfg("dpsub", 1, memo_kind=GQ_MEMO_DENSE))
    hashed, hst = gpu_bits(g, cfg("dpsub", 1, memo_kind=GQ_MEMO_HASH))
    ranked, rst = gpu_bits(g, cfg("dpsub", 1, memo_kind=GQ_MEMO_RANK))
    assert_same_plan(dense, want, "snowflake-30 dense vs oracle")
    assert_same_plan(hashed, want, "snowflake-30 hash vs oracle")
    assert_same_plan(ranked, want, "snowflake-30 rank vs oracle")
    assert (dst.pairs, dst.sets) == (rec["pairs"], rec["sets"])
    assert hst.pairs == dst.pairs and hst.sets == dst.sets
    assert rst.pairs == dst.pairs and rst.sets == dst.sets
    assert hst.memo_bytes * 50 < dst.memo_bytes          # 2^30 dense slots vs ~4M connected sets
    assert rst.memo_bytes < hst.memo_bytes                # no keys, one slot per connected set


@pytest.mark.parametrize("kind", [0, GQ_MEMO_HASH], ids=["auto", "hash"])
def test_server_with_tree_memo_requests(kind):
    """Requests above the dense limit share engine steps with dense ones
    (auto: rank memo; hash: open addressing)."""
    import gpuqo_amd
    direct = cfg("dpsub", 1, memo_kind=kind)
    reqs = [qgen.snowflake(26, 0x26), qgen.snowflake(12, 0x12), qgen.snowflake(24, 0x24), qgen.star(14, 0x14),
            qgen.snowflake(28, 0x28, depth=4, fanout=5), qgen.snowflake(10, 0x10)]
    with gpuqo_amd.Server(direct, 2 << 30, max_inflight=4) as srv:
        for i, g in enumerate(reqs):
            srv.submit(g, i)
        got = {}
        while len(got) < len(reqs):
            r = srv.poll(wait_ms=60000)
            assert r is not None
            tag, st, nodes, stats = r
            assert st == 0, f"request {tag}: status {st}"
            got[tag] = (nodes, stats)
    from test_golden_large import tree_fixture
    fixtures = {0: "snowflake26-mpdp", 2: "snowflake24-mpdp", 4: "snowflake28d4f5-mpdp"}
    for i, g in enumerate(reqs):
        alone, ast = gpu_bits(g, direct)
        assert_same_plan(gpuqo_amd.plan_bits(got[i][0]), alone, f"request {i} ({g.shape}{g.n})")
        assert got[i][1].memo_kind == ast.memo_kind
        if i in fixtures:                         # the oracle's plans of the large ones
            fg, want, rec = tree_fixture(fixtures[i])
            assert fg.n == g.n
            assert_same_plan(alone, want, f"request {i} ({g.shape}{g.n}) vs oracle")
            assert got[i][1].pairs == rec["pairs"]


@pytest.mark.parametrize("idp,k,n", [("idp2", 8, 20), ("union", 25, 40), ("idp1", 6, 16)])
def test_rank_memo_under_idp_drivers(idp, k, n):
    """The rank memo requested for every DP of the IDP1, IDP2 and UnionDP
    drivers (whose sub-DPs are small trees, relabelled and contracted)."""
    g = qgen.snowflake(n, 0x3000 + n, depth=4, fanout=5) if n > 30 else qgen.snowflake(n, 0x3000 + n)
    c = cfg("dpsub", 1, idp_type=idp, idp_n_iters=k, memo_kind=GQ_MEMO_RANK)
    want, _ = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, f"{idp} k={k} snowflake-{n} rank memo")


def test_server_with_rank_memo_everywhere():
    """Every tree DP of a mixed request set on the rank memo, on the engine."""
    import gpuqo_amd
    direct = cfg("dpsub", 1, memo_kind=GQ_MEMO_RANK)
    union = cfg("dpsub", 1, idp_type="union", idp_n_iters=25, memo_kind=GQ_MEMO_RANK)
    reqs = [(qgen.snowflake(18, 0x3101), direct), (qgen.star(12, 0x3102), direct), (qgen.cycle(12, 0x3103), direct),
            (qgen.snowflake(70, 0x3104, depth=4, fanout=5), union), (qgen.chain(16, 0x3105), direct)]
    with gpuqo_amd.Server(direct, 2 << 30, max_inflight=4) as srv:
        for i, (g, c) in enumerate(reqs):
            srv.submit(g, i, c)
        got = {}
        while len(got) < len(reqs):
            r = srv.poll(wait_ms=60000)
            assert r is not None
            assert r[1] == 0, f"request {r[0]}: status {r[1]}"
            got[r[0]] = r[2]
    for i, (g, c) in enumerate(reqs):
        want, _ = oracle_bits(g, c)
        assert_same_plan(gpuqo_amd.plan_bits(got[i]), want, f"request {i} ({g.shape}{g.n}) vs oracle")
