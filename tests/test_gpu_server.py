"""The continuous-batching planning server: every request's plan equals the
same query planned alone (gq_plan) and the oracle's, whatever else shares the
device with it."""
import time

import pytest

from helpers import assert_same_plan, cfg, dpsub_sparse_graphs, gpu_bits, oracle_bits
from gpuqo_amd import qgen

pytestmark = pytest.mark.gpu


def _mix():
    direct = cfg("dpsub", 1)
    union = cfg("dpsub", 1, idp_type="union", idp_n_iters=25)
    reqs = [(qgen.snowflake(n, 5000 + n), direct) for n in (10, 12, 14, 16, 18, 20)]
    reqs += [(qgen.snowflake(70, 77, depth=4, fanout=5), union), (qgen.cycle(16, 99), direct),
             (qgen.star(13, 98), direct), (qgen.snowflake(64, 78, depth=4, fanout=5), union),
             (qgen.chain(15, 97), direct), (qgen.snowflake(22, 96), direct)]
    return direct, reqs


def _run(srv, reqs, gap_s):
    for i, (g, c) in enumerate(reqs):
        srv.submit(g, i, c)
        if gap_s:
            time.sleep(gap_s)                     # later requests arrive while earlier ones run
    got = {}
    while len(got) < len(reqs):
        r = srv.poll(wait_ms=60000)
        assert r is not None, f"server returned nothing after {len(got)} of {len(reqs)}"
        tag, st, nodes, stats = r
        assert st == 0, f"request {tag}: status {st}"
        assert tag not in got
        assert stats.wall_ms > 0.0
        got[tag] = nodes
    return got


def test_server_mix_equals_alone_and_oracle():
    import gpuqo_amd
    direct, reqs = _mix()
    with gpuqo_amd.Server(direct, 4 << 30, max_inflight=6) as srv:
        got = _run(srv, reqs, 0.002)
        assert srv.poll(wait_ms=0) is None
    for i, (g, c) in enumerate(reqs):
        plan = gpuqo_amd.plan_bits(got[i])
        alone, _ = gpu_bits(g, c)
        assert_same_plan(plan, alone, f"request {i} ({g.shape}{g.n}) vs alone")
        want, _ = oracle_bits(g, c)
        assert_same_plan(plan, want, f"request {i} ({g.shape}{g.n}) vs oracle")


def test_server_dpsub_sparse_cyclic_requests():
    """Plain DPsub requests on sparse cyclic graphs (their splits found
    through the blocks) on the step engine equal the oracle's plans."""
    import gpuqo_amd
    c = cfg("dpsub", 0)
    reqs = [(g, c) for _, g in dpsub_sparse_graphs()]
    with gpuqo_amd.Server(c, 4 << 30, max_inflight=4) as srv:
        got = _run(srv, reqs, 0.0)
    for i, (g, _) in enumerate(reqs):
        want, _ = oracle_bits(g, c)
        assert_same_plan(gpuqo_amd.plan_bits(got[i]), want, f"served dpsub {g.shape}{g.n}")


def test_server_tight_budget_admits_one_at_a_time():
    """A budget that holds one 20-relation memo: requests are admitted as
    memory frees up, and all finish with the same plans."""
    import gpuqo_amd
    direct = cfg("dpsub", 1)
    reqs = [(qgen.snowflake(n, 6100 + n), direct) for n in (20, 18, 19, 20, 16)]
    # one 20-relation DP takes 2^20 x 36 B plus its query blob: about 38 MB
    with gpuqo_amd.Server(direct, 40 << 20, max_inflight=4) as srv:
        got = _run(srv, reqs, 0.0)
    for i, (g, c) in enumerate(reqs):
        alone, _ = gpu_bits(g, c)
        assert_same_plan(gpuqo_amd.plan_bits(got[i]), alone, f"request {i}")


def test_server_rejects_a_dp_larger_than_its_budget():
    import gpuqo_amd
    direct = cfg("dpsub", 1)
    with gpuqo_amd.Server(direct, 8 << 20, max_inflight=2) as srv:
        srv.submit(qgen.snowflake(20, 6200), 1)
        srv.submit(qgen.snowflake(8, 6201), 2)
        res = {}
        for _ in range(2):
            r = srv.poll(wait_ms=60000)
            assert r is not None
            res[r[0]] = r[1]
    assert res[1] == 3            # GQ_ERR_NOMEM: 2^20 memo slots do not fit 8 MiB
    assert res[2] == 0


def test_server_burst_of_mixed_requests():
    """A burst of 48 requests submitted at once: dense, rank-memo and
    general-graph DPs and UnionDP requests share steps, arena slices are
    freed and reused while later steps are in flight, and every plan equals
    the same request planned alone."""
    import gpuqo_amd
    direct = cfg("dpsub", 1)
    union = cfg("dpsub", 1, idp_type="union", idp_n_iters=25)
    reqs = []
    for i in range(48):
        r = i % 6
        if r == 0:
            reqs.append((qgen.snowflake(10 + i % 13, 7000 + i), direct))
        elif r == 1:
            reqs.append((qgen.snowflake(25 + i % 3, 7100 + i), direct))       # rank memo
        elif r == 2:
            reqs.append((qgen.cycle(9 + i % 8, 7200 + i), direct))
        elif r == 3:
            reqs.append((qgen.star(8 + i % 7, 7300 + i), direct))
        elif r == 4:
            reqs.append((qgen.snowflake(60 + i, 7400 + i, depth=4, fanout=5), union))
        else:
            reqs.append((qgen.chain(6 + i % 10, 7500 + i), direct))
    with gpuqo_amd.Server(direct, 8 << 30, max_inflight=32) as srv:
        got = _run(srv, reqs, 0.0)
    for i, (g, c) in enumerate(reqs):
        alone, _ = gpu_bits(g, c)
        assert_same_plan(gpuqo_amd.plan_bits(got[i]), alone, f"request {i} ({g.shape}{g.n})")


def test_server_spanning_tree_requests_beside_others():
    """Requests planned over their minimum spanning tree (trees whose classes
    are not all tree edges) share engine steps with ordinary tree and cyclic
    requests; each equals the oracle."""
    import gpuqo_amd
    span = cfg("dpsub", 1, spanning_tree=1)
    span_dpsub = cfg("dpsub", 0, spanning_tree=1)
    direct = cfg("dpsub", 1)
    reqs = [(qgen.cycle(n, 0x5b00 + n), span) for n in (10, 14, 18)]
    reqs += [(qgen.chorded_cycle(20, 5, 0x5b20), span), (qgen.clique(11, 0x5b11), span_dpsub),
             (qgen.chorded_cycle(40, 7, 0x5b40), cfg("dpsub", 1, spanning_tree=1, idp_type="union", idp_n_iters=25)),
             (qgen.snowflake(18, 0x5b18), direct), (qgen.cycle(12, 0x5b12), direct), (qgen.star(12, 0x5b0c), direct)]
    with gpuqo_amd.Server(direct, 4 << 30, max_inflight=8) as srv:
        got = _run(srv, reqs, 0.0)
    for i, (g, c) in enumerate(reqs):
        want, _ = oracle_bits(g, c)
        assert_same_plan(gpuqo_amd.plan_bits(got[i]), want, f"request {i} ({g.shape}{g.n})")
