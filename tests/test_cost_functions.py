"""The reference's build-time cost functions (Makefile cost_function =
postgres | simple | cout, gpuqo_cost.cuh:25-38), selected here by
gq_config.cost_function.

CPU tests pin the oracle: hand-derived known answers for cost_baserel and
calc_join_cost under simple (gpuqo_cost_simple.cuh) and C_out
(gpuqo_cost_cout.cuh), and the DP optimum against an enumeration of every
plan of small queries whose row counts are exact (powers of two), so that a
set's rows do not depend on the plan and the cheapest plan's total is the
minimum over all plans.  GPU tests compare the HIP path with the oracle bit
for bit.
"""
import itertools

import numpy as np
import pytest

import gpuqo_amd
import oracle
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan, cfg, gpu_bits, oracle_bits, small_graphs

JOIN, BASE = oracle.PROBE_JOIN, oracle.PROBE_BASEREL
INF = float("inf")


def kcfg(cost_function, **kw):
    return abi.default_config(algorithm="cpu_dpsub", cost_function=cost_function, **kw)


def two_rel(index_total=None):
    """rel0 (outer) and rel1 (inner) joined by one generic class; with
    index_total, rel1's member of the class has an available index of that
    total cost."""
    g = qgen.QueryGraph(2, "kat")
    g.rows = [10.0, 20.0]
    g.add_generic(0, 1, 0.125)
    g.build()
    g.base[0]["tuples"] = 1000.0
    g.base[1]["tuples"] = 40.0
    if index_total is not None:
        slot = g.ec_offsets[0] + 1
        for k, x in dict(index_available=1, index_unique=0, index_rows=3.0, index_startup=0.5,
                         index_total=index_total).items():
            g.ec_vars[slot][k] = x
    return g


# ------------------------------------------------------------ cost_baserel

@pytest.mark.parametrize("fn,composite,expect", [
    # postgres: the relation's own cost (cost_postgres.cuh:843-847)
    ("postgres", 0, (3.0, 77.5)),
    ("postgres", 1, (3.0, 77.5)),
    # simple: startup 0, total BASEREL_COEFF * tuples = 0.2f * 1000 = 200
    # (0.2f * 1000 = 200.0000030 rounds to 200.0f), or the composite's total
    ("simple", 0, (0.0, 200.0)),
    ("simple", 1, (0.0, 77.5)),
    # C_out: startup 0, total 0, or the composite's total
    ("cout", 0, (0.0, 0.0)),
    ("cout", 1, (0.0, 77.5)),
])
def test_cost_baserel(fn, composite, expect):
    """Relation 0 has rows 10 and tuples 1000 (tuples, not rows, scale the
    simple cost), cost {3, 77.5}."""
    g = two_rel()
    g.base[0]["startup"], g.base[0]["total"] = 3.0, 77.5
    assert oracle.cost_probe(g, kcfg(fn), BASE, [1], [2], composite_mask=composite) == expect


# ------------------------------------------------------------ calc_join_cost

L10 = (10.0, 3.0, 5.0, 8.0)      # outer: rows 10, startup 3, total 5
R20 = (20.0, 1.0, 7.0, 4.0)      # inner: rows 20, startup 1, total 7
EXTRA = (0, 0, 50.0, 0, 0)       # join rows 50


@pytest.mark.parametrize("fn,hashjoin,index_total,composite,expect", [
    # simple: nestloop 5 + 10 * 7 = 75; hashjoin 1 * 50 + 7 + 5 = 62;
    # index nested loop over the inner's index: 5 + 10 * 2 = 25
    ("simple", 1, None, 0, (0.0, 62.0)),
    ("simple", 0, None, 0, (0.0, 75.0)),
    ("simple", 1, 2.0, 0, (0.0, 25.0)),
    ("simple", 1, 6.0, 0, (0.0, 62.0)),     # 5 + 60 = 65 is not below 62
    ("simple", 0, 6.0, 0, (0.0, 65.0)),     # ... but below the nested loop's 75
    ("simple", 1, 2.0, 2, (0.0, 62.0)),     # a composite inner has no index path
    # C_out: nestloop +inf; hashjoin 50 + 7 + 5 = 62; index path +inf
    ("cout", 1, None, 0, (0.0, 62.0)),
    ("cout", 1, 2.0, 0, (0.0, 62.0)),
    ("cout", 0, 2.0, 0, (0.0, INF)),
])
def test_calc_join_cost(fn, hashjoin, index_total, composite, expect):
    g = two_rel(index_total)
    out = oracle.cost_probe(g, kcfg(fn, enable_hashjoin=hashjoin), JOIN, [1], [2], L10, R20, EXTRA,
                            composite_mask=composite)
    assert out == expect


def test_postgres_is_the_default():
    """The default config selects the postgres cost function: calc_join_cost
    then gives a startup cost (the simple and C_out ones never do)."""
    g = two_rel()
    assert abi.default_config().cost_function == abi.COST_FUNCTION["postgres"] == 0
    out = oracle.cost_probe(g, kcfg("postgres"), JOIN, [1], [2], L10, R20, EXTRA)
    assert out[0] > 0.0


# ------------------------------------------------------------ DP optimum against every plan

def exact_graph(n, edges, seed):
    """Rows 2^6..2^10 and selectivities 2^-3..2^-1: every join's rows are an
    exact power of two above 1, the same whatever the plan."""
    rng = np.random.default_rng(seed)
    g = qgen.QueryGraph(n, "exact")
    g.rows = [float(2 ** int(rng.integers(6, 11))) for _ in range(n)]
    for (u, v) in edges:
        g.add_generic(u, v, float(2.0 ** -int(rng.integers(1, 4))))
    g.build()
    return g


def all_plan_totals(g, fn):
    """(rows, total) of every plan of every connected set, built bottom-up
    from explicit trees: a join of disjoint connected sets with a class
    between them, either side outer.  float32 throughout, as the cost code."""
    n = len(g.rows)
    f = np.float32
    sel = {}
    for (u, v, _kind, s) in g.edges:
        sel[frozenset((u, v))] = f(s)
    adj = {i: set() for i in range(n)}
    for k in sel:
        a, b = tuple(k)
        adj[a].add(b)
        adj[b].add(a)

    def connected(S):
        S = set(S)
        seen, stack = set(), [next(iter(S))]
        while stack:
            x = stack.pop()
            if x in seen:
                continue
            seen.add(x)
            stack += [y for y in adj[x] if y in S and y not in seen]
        return seen == S

    def crossing(A, B):
        """Product of the selectivities of the classes between A and B
        (None: no class, a cross product, which no DP considers)."""
        out, any_edge = f(1.0), False
        for k, s in sel.items():
            a, b = tuple(k)
            if (a in A and b in B) or (a in B and b in A):
                out, any_edge = f(out * s), True
        return out if any_edge else None

    memo = {}

    def plans(S):
        if S in memo:
            return memo[S]
        res = []
        if len(S) == 1:
            (i,) = S
            tuples = f(g.base[i]["tuples"])
            base = {"simple": f(f(0.2) * tuples), "cout": f(0.0)}[fn]
            res.append((f(g.rows[i]), base))
        else:
            items = sorted(S)
            for r in range(1, len(items)):
                for A in itertools.combinations(items, r):
                    A = frozenset(A)
                    B = S - A
                    if not connected(A) or not connected(B):
                        continue
                    s = crossing(A, B)
                    if s is None:
                        continue
                    for (lr, lt) in plans(A):          # A outer, B inner
                        for (rr, rt) in plans(B):
                            rows = f(f(s * lr) * rr)
                            rows = f(np.rint(rows)) if rows > 1 else f(1.0)
                            nl = f(lt + f(lr * rt)) if fn == "simple" else f(INF)
                            hj = f(f(rows + rt) + lt)
                            res.append((rows, hj if hj < nl else nl))
        memo[S] = res
        return res

    return plans(frozenset(range(n)))


SHAPES = [
    ("chain5", 5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
    ("star5", 5, [(0, 1), (0, 2), (0, 3), (0, 4)]),
    ("cycle5", 5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]),
    ("clique4", 4, [(a, b) for a in range(4) for b in range(a + 1, 4)]),
]


@pytest.mark.parametrize("fn", ["simple", "cout"])
@pytest.mark.parametrize("name,n,edges", SHAPES)
@pytest.mark.parametrize("algo,bicc", [("cpu_dpsub", 0), ("cpu_dpsub_bicc", 1), ("cpu_dpsize", 0)])
def test_oracle_optimum_is_the_cheapest_plan(fn, name, n, edges, algo, bicc):
    g = exact_graph(n, edges, 7 + n)
    totals = all_plan_totals(g, fn)
    best = min(t for (_, t) in totals)
    c = abi.default_config(algorithm=algo, dpsub_bicc=bicc, cost_function=fn)
    st, nodes, _ = oracle.plan(g, c)
    assert st == 0
    assert np.float32(nodes[0][3]) == best, (name, nodes[0], best)
    assert nodes[0][2] == 0.0                 # simple and C_out: startup 0


def test_cost_function_out_of_range_is_refused():
    """gq_plan rejects a cost_function outside 0..2 before any device work."""
    g = qgen.star(4, 3)
    c = cfg("dpsub", cost_function=3)
    with pytest.raises(gpuqo_amd.GpuqoError) as e:
        gpuqo_amd.plan(g, c)
    assert e.value.status == abi.GQ_ERR_INVALID
    assert "cost_function" in str(e.value)


# ------------------------------------------------------------ GPU parity

GRAPHS = [(n, g) for n, g in small_graphs() if g.n >= 3]


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["simple", "cout"])
@pytest.mark.parametrize("algo,bicc", [("dpsub", 0), ("dpsub", 1), ("dpsize", 0)])
def test_gpu_small_graphs(fn, algo, bicc):
    oalgo = {"dpsub": "cpu_dpsub_bicc" if bicc else "cpu_dpsub", "dpsize": "cpu_dpsize"}[algo]
    for name, g in GRAPHS:
        a, _ = gpu_bits(g, cfg(algo, bicc, cost_function=fn))
        b, _ = oracle_bits(g, cfg(oalgo, bicc, cost_function=fn))
        assert_same_plan(a, b, f"{name} {algo} bicc={bicc} {fn}")


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["simple", "cout"])
@pytest.mark.parametrize("name,g,bicc", [
    ("snowflake20", qgen.query_snowflake20(), 1),
    ("star12", qgen.query_star12(), 0),
    ("clique12", qgen.clique(12, 5), 0),
    ("cycle18", qgen.cycle(18, 9), 1),
])
def test_gpu_headline_shapes(fn, name, g, bicc):
    a, sa = gpu_bits(g, cfg("dpsub", bicc, cost_function=fn))
    b, sb = oracle_bits(g, cfg("cpu_dpsub_bicc" if bicc else "cpu_dpsub", bicc, cost_function=fn))
    assert_same_plan(a, b, f"{name} {fn}")
    assert sa.pairs == sb.pairs


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["simple", "cout"])
@pytest.mark.parametrize("memo", ["rank", "hash"])
def test_gpu_tree_memos(fn, memo):
    """A 24-relation snowflake through the rank and open-addressing memos."""
    g = qgen.snowflake(24, 17)
    a, sa = gpu_bits(g, cfg("dpsub", 1, cost_function=fn, memo_kind=memo))
    b, _ = oracle_bits(g, cfg("cpu_dpsub_bicc", 1, cost_function=fn))
    assert_same_plan(a, b, f"snowflake24 {memo} {fn}")
    assert sa.memo_kind == abi.MEMO[memo]


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["simple", "cout"])
@pytest.mark.parametrize("idp,k", [("idp2", 8), ("union", 10)])
def test_gpu_idp(fn, idp, k):
    """IDP2 and UnionDP over a 40-relation snowflake: composite relations
    enter the later DPs with their sub-plans' totals (startup 0)."""
    g = qgen.snowflake(40, 23)
    a, _ = gpu_bits(g, cfg("dpsub", 1, cost_function=fn, idp_type=idp, idp_n_iters=k))
    b, _ = oracle_bits(g, cfg("cpu_dpsub_bicc", 1, cost_function=fn, idp_type=idp, idp_n_iters=k))
    assert_same_plan(a, b, f"snowflake40 {idp} k={k} {fn}")


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["simple", "cout"])
def test_gpu_wide(fn):
    """A 40-relation chain in the wide (64-bit set) DP."""
    g = qgen.chain(40, 29)
    a, _ = gpu_bits(g, cfg("dpsub", 1, cost_function=fn))
    st, nodes, _ = oracle.plan_connected(g, cfg("cpu_dpsub_bicc", 1, cost_function=fn))
    assert st == 0
    assert_same_plan(a, gpuqo_amd.plan_bits(nodes), f"chain40 {fn}")


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["simple", "cout"])
def test_gpu_batch_and_server(fn):
    """The batched form and the server plan as gq_plan does."""
    qs = [qgen.snowflake(n, 100 + n) for n in range(8, 20)] + [qgen.cycle(n, 200 + n) for n in range(6, 14)]
    c = cfg("dpsub", 1, cost_function=fn)
    batch = gpuqo_amd.plan_batch(qs, c)
    for i, g in enumerate(qs):
        b, _ = oracle_bits(g, cfg("cpu_dpsub_bicc", 1, cost_function=fn))
        assert_same_plan(gpuqo_amd.plan_bits(batch[i][0]), b, f"batch {i} {fn}")
    with gpuqo_amd.Server(c, 1 << 30, max_inflight=8) as srv:
        for i, g in enumerate(qs):
            srv.submit(g, i)
        got = {}
        for _ in qs:
            r = srv.poll(60000)
            assert r is not None and r[1] == 0
            got[r[0]] = gpuqo_amd.plan_bits(r[2])
    for i, g in enumerate(qs):
        assert got[i] == gpuqo_amd.plan_bits(batch[i][0]), f"server {i} {fn}"
