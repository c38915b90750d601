"""The open-addressing memo (tree graphs): the same plans as the dense memo
and the oracle, with level lists generated by expansion."""
import pytest

from helpers import assert_same_plan, cfg, gpu_bits, oracle_bits
from gpuqo_amd import qgen

pytestmark = pytest.mark.gpu

GQ_MEMO_DENSE, GQ_MEMO_HASH = 1, 2


def _trees():
    return [("snowflake20", qgen.config2(), 1), ("star12", qgen.config1(), 0), ("chain15", qgen.chain(15, 7), 1),
            ("snow16-dpsub", qgen.snowflake(16, 33), 0), ("star9", qgen.star(9, 5), 1),
            ("chain2", qgen.chain(2, 1), 1), ("snow3", qgen.snowflake(3, 2), 0),
            ("snow18-d4", qgen.snowflake(18, 44, depth=4, fanout=5), 1)]


@pytest.mark.parametrize("name,g,bicc", _trees(), ids=lambda x: x if isinstance(x, str) else "")
def test_hash_memo_matches_oracle(name, g, bicc):
    c = cfg("dpsub", bicc, memo_kind=GQ_MEMO_HASH)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, name)
    assert gst.memo_kind == GQ_MEMO_HASH
    assert gst.pairs == ost.pairs
    assert gst.sets == ost.sets


def test_auto_memo_switches_above_the_memo_limit():
    g = qgen.snowflake(22, 0x22)
    want, _ = oracle_bits(g, cfg("dpsub", 1))
    got, gst = gpu_bits(g, cfg("dpsub", 1, max_memo_size_mb=16))      # dense: 36 B x 2^22 = 151 MB
    assert gst.memo_kind == GQ_MEMO_HASH
    assert_same_plan(got, want, "auto -> open addressing")
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


def test_hash_equals_dense_at_30_relations():
    g = qgen.snowflake(30, 0x30)
    dense, dst = gpu_bits(g, cfg("dpsub", 1, memo_kind=GQ_MEMO_DENSE))
    hashed, hst = gpu_bits(g, cfg("dpsub", 1, memo_kind=GQ_MEMO_HASH))
    assert_same_plan(hashed, dense, "snowflake-30")
    assert hst.pairs == dst.pairs and hst.sets == dst.sets
    assert hst.memo_bytes * 50 < dst.memo_bytes          # 2^30 dense slots vs ~4M connected sets


def test_server_with_open_addressing_requests():
    import gpuqo_amd
    direct = cfg("dpsub", 1)
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
    for i, g in enumerate(reqs):
        alone, ast = gpu_bits(g, direct)
        assert_same_plan(gpuqo_amd.plan_bits(got[i][0]), alone, f"request {i} ({g.shape}{g.n})")
        assert got[i][1].memo_kind == ast.memo_kind
