"""Pins the CPU oracle (oracle/gq_oracle.cpp) to facts fixed independently of
its code: closed-form pair counts, brute-force enumeration, textbook graph
algorithms, hand-derived cost-model answers and an independent restatement
of the DP recurrence.  The reference ships no golden vectors for this path
and cannot be compiled here, so these are the oracle's anchors."""
import copy
import itertools
from collections import deque

import numpy as np
import pytest

import gpuqo_amd
import oracle
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan, oracle_bits, small_graphs


def cfg(algo, **kw):
    return abi.default_config(algorithm=algo, **kw)


def connected(mask, adj):
    if mask == 0:
        return False
    start = mask & -mask
    seen, stack = start, [start.bit_length() - 1]
    while stack:
        u = stack.pop()
        nb = adj[u] & mask & ~seen
        while nb:
            b = nb & -nb
            nb ^= b
            seen |= b
            stack.append(b.bit_length() - 1)
    return seen == mask


def brute_pairs_and_sets(g):
    adj = g.adjacency()
    n = g.n
    P = S = 0
    for s in range(1, 1 << n):
        if bin(s).count("1") < 2 or not connected(s, adj):
            continue
        S += 1
        sub = (s - 1) & s
        while sub:
            if connected(sub, adj) and connected(s ^ sub, adj):
                P += 1
            sub = (sub - 1) & s
    return P, S


# ---------------------------------------------------------------- counts

@pytest.mark.parametrize("n", [3, 5, 8, 11])
def test_pair_counts_closed_forms(n):
    cases = [
        (qgen.star(n, 1), (n - 1) * 2 ** (n - 1)),
        (qgen.clique(n, 2), 3 ** n - 2 ** (n + 1) + 1),
        (qgen.chain(n, 3), (n ** 3 - n) // 3),
        (qgen.cycle(n, 4), n ** 3 - 2 * n ** 2 + n),
    ]
    for g, want in cases:
        for algo in ("cpu_dpsub", "cpu_dpsub_bicc", "cpu_dpsize"):
            st, nodes, stats = oracle.plan(g, cfg(algo))
            assert st == 0
            assert stats.pairs == want, (g.shape, n, algo)
            assert len(nodes) == 2 * n - 1


@pytest.mark.parametrize("n,seed", [(7, 1), (10, 2), (12, 3), (14, 4)])
def test_tree_pairs_and_sets_brute_force(n, seed):
    g = qgen.snowflake(n, seed)
    P, S = brute_pairs_and_sets(g)
    for algo in ("cpu_dpsub", "cpu_dpsub_bicc"):
        st, _, stats = oracle.plan(g, cfg(algo))
        assert st == 0 and stats.pairs == P and stats.sets == S
    # trees: every connected set of size s has exactly 2(s-1) splits
    adj = g.adjacency()
    tot = sum(2 * (bin(s).count("1") - 1) for s in range(1, 1 << n)
              if bin(s).count("1") > 1 and connected(s, adj))
    assert tot == P


# ---------------------------------------------------------------- graph algorithms

def textbook_bfs(g):
    adj = g.adjacency()
    order, seen, q = [], {0}, deque([0])
    while q:
        u = q.popleft()
        order.append(u)
        for v in range(g.n):
            if adj[u] >> v & 1 and v not in seen:
                seen.add(v)
                q.append(v)
    return order


@pytest.mark.parametrize("name,g", small_graphs(), ids=lambda x: x if isinstance(x, str) else "")
def test_bfs_order(name, g):
    assert oracle.bfs_order(g) == textbook_bfs(g)


def tarjan_blocks(mask, adj):
    """Recursive Hopcroft-Tarjan with an edge stack, neighbours ascending."""
    disc, low, out, es = {}, {}, [], []
    t = [0]

    def dfs(u, parent):
        disc[u] = low[u] = t[0]
        t[0] += 1
        for v in range(32):
            if not (adj[u] >> v & 1 and mask >> v & 1):
                continue
            if v not in disc:
                es.append((u, v))
                dfs(v, u)
                low[u] = min(low[u], low[v])
                if low[v] >= disc[u]:
                    blk = 0
                    while True:
                        a, b = es.pop()
                        blk |= 1 << a | 1 << b
                        if (a, b) == (u, v):
                            break
                    out.append(blk)
            elif v != parent and disc[v] < disc[u]:
                es.append((u, v))
                low[u] = min(low[u], disc[v])

    dfs((mask & -mask).bit_length() - 1, -1)
    return out


def random_graph(n, p, seed):
    import random
    r = random.Random(seed)
    g = qgen.QueryGraph(n, "random")
    for i in range(n):
        g.rows[i] = float(r.randint(10, 1000))
    for i in range(1, n):                      # spanning tree keeps it connected
        g.add_generic(r.randrange(i), i, 0.01)
    have = {(min(a, b), max(a, b)) for a, b, _, _ in g.edges}
    for i in range(n):
        for j in range(i + 1, n):
            if (i, j) not in have and r.random() < p:
                g.add_generic(i, j, 0.05)
    return g.build()


@pytest.mark.parametrize("seed", range(6))
def test_block_order_matches_recursive_tarjan(seed):
    g = random_graph(9, 0.25, seed)
    adj = g.adjacency()
    for s in range(1, 1 << g.n):
        if bin(s).count("1") >= 2 and connected(s, adj):
            assert oracle.blocks(g, s) == tarjan_blocks(s, adj), hex(s)


@pytest.mark.parametrize("seed", range(4))
def test_bicc_splits_are_exactly_the_csg_cmp_pairs(seed):
    """Order-free: MPDP must cost the same pairs as plain DPsub (and find the
    same optimum when no two splits tie)."""
    g = random_graph(9, 0.3, 100 + seed)
    P, S = brute_pairs_and_sets(g)
    st, a, sa = oracle.plan(g, cfg("cpu_dpsub"))
    st2, b, sb = oracle.plan(g, cfg("cpu_dpsub_bicc"))
    assert sa.pairs == sb.pairs == P and sa.sets == sb.sets == S
    assert a[0][3] == b[0][3]


def test_dpsize_level1_order_is_a_permutation():
    for n in (2, 5, 12, 31):
        o = oracle.dpsize_level1_order(n)
        assert sorted(o) == list(range(n))


# ---------------------------------------------------------------- cost model

def test_hash_table_size_known_answers():
    # 100 tuples of width 8: 1024 buckets, 1 batch, 83886 // 124 skew MCVs
    assert oracle.hash_table_size(100.0, 8) == (1024, 1, 676)
    # 1e6 tuples: 40 MB > work_mem -> lbuckets = 2^ceil_log2(4110418 // 48) = 131072,
    # batches = 2^ceil_log2(ceil(4e7 / (4110418 - 1048576))) = 16
    assert oracle.hash_table_size(1.0e6, 8) == (131072, 16, 676)
    # non-positive input rows are forced to 1000
    assert oracle.hash_table_size(0.0, 8) == oracle.hash_table_size(1000.0, 8)
    # batches are capped at max_pointers = 524288
    assert oracle.hash_table_size(1.0e12, 8)[1] == 524288


def two_rel(rows0, rows1, sel, fk=False, idx=False):
    g = qgen.QueryGraph(2, "kat")
    g.rows = [rows0, rows1]
    if fk:
        g.add_pkfk(0, 1)
    else:
        g.add_generic(0, 1, sel)
    g.build()
    if not idx:
        g.ec_vars["index_available"] = 0
        g.ec_vars["index_unique"] = 0
    return g


def test_nestloop_known_answer_dyadic():
    """Only nested loop enabled, dyadic parameters: every operation is exact.
    outer {rows 8, cost 0..64}, inner {rows 4, cost 0..16}, sel 1/2:
      rows = 8*4/2 = 16
      run  = 64 + 16 + (8-1)*16 = 192, cpu = (1/4 + 1/8 per qual) * 32 = 12
      total = 204"""
    g = two_rel(8.0, 4.0, 0.5)
    c = cfg("cpu_dpsub", cpu_tuple_cost=0.25, cpu_operator_cost=0.125, enable_hashjoin=0)
    out = oracle.join_cost(g, c, [1], [2], (8.0, 0.0, 64.0, 8.0), (4.0, 0.0, 16.0, 4.0))
    assert out == (16.0, 0.0, 204.0, 12.0)


def test_nestloop_startup_terms():
    """Inner startup 2 is paid once up front and again per extra outer row:
    startup = 2 + 1 = 3; run = (10-1) + 7*2 + 14 + 7*14 = 135;
    cpu = (0.25+0.125)*8*8 = 24; total = 3 + 135 + 24 = 162."""
    g = two_rel(8.0, 8.0, 0.25)
    c = cfg("cpu_dpsub", cpu_tuple_cost=0.25, cpu_operator_cost=0.125, enable_hashjoin=0)
    out = oracle.join_cost(g, c, [1], [2], (8.0, 1.0, 10.0, 8.0), (8.0, 2.0, 16.0, 8.0))
    assert out == (16.0, 3.0, 162.0, 16.0)


def test_disable_cost_and_hashjoin_choice():
    """With nested loop disabled its cost carries disable_cost (2^20 here), so
    the hash join must win and its cost cannot contain it."""
    g = two_rel(8.0, 4.0, 0.5)
    c = cfg("cpu_dpsub", cpu_tuple_cost=0.25, cpu_operator_cost=0.125, enable_nestloop=0,
            disable_cost=float(2 ** 20))
    rows, startup, total, width = oracle.join_cost(g, c, [1], [2], (8.0, 0.0, 64.0, 8.0),
                                                   (4.0, 0.0, 16.0, 4.0))
    assert rows == 16.0 and total < 2 ** 20
    # hash join: startup = inner total 16 + (1/8*1 + 1/4)*4 = 17.5
    assert startup == 17.5


def test_fk_selectivity_is_double_division_narrowed():
    g = two_rel(1000.0, 3.0, 0.0, fk=True)
    c = cfg("cpu_dpsub", enable_hashjoin=0)
    rows = oracle.join_cost(g, c, [1], [2], (1000.0, 0.0, 10.0, 8.0), (3.0, 0.0, 1.0, 4.0))[0]
    import numpy as np
    sel = np.float32(1.0 / 3.0)
    assert rows == float(np.round(np.float32(np.float32(sel * np.float32(1000.0)) * np.float32(3.0))))


def test_rows_clamp_and_rounding():
    g = two_rel(10.0, 10.0, 1e-6)
    c = cfg("cpu_dpsub")
    assert oracle.join_cost(g, c, [1], [2], (10.0, 0.0, 1.0, 4.0), (10.0, 0.0, 1.0, 4.0))[0] == 1.0
    g = two_rel(3.0, 5.0, 0.1)          # 1.5000000x -> 2 (round half away from zero)
    r = oracle.join_cost(g, c, [1], [2], (3.0, 0.0, 1.0, 4.0), (5.0, 0.0, 1.0, 4.0))[0]
    assert r == 2.0


def test_commuted_join_costs_differ():
    """Outer/inner roles matter: a dropped or swapped operand changes the cost."""
    g = qgen.clique(3, 9)
    c = cfg("cpu_dpsub")
    a = oracle.join_cost(g, c, [1], [2], (100.0, 0.0, 50.0, 4.0), (10.0, 0.0, 5.0, 4.0))
    b = oracle.join_cost(g, c, [2], [1], (10.0, 0.0, 5.0, 4.0), (100.0, 0.0, 50.0, 4.0))
    assert a[0] == b[0] and a[2] != b[2]


# ---------------------------------------------------------------- DP recurrence

def python_dpsub(g, c):
    """Independent restatement of the DPsub recurrence: sets in increasing
    integer order, splits l ascending, first found kept unless a later split
    is strictly cheaper; joins costed by the oracle's cost model."""
    adj = g.adjacency()
    n = g.n
    nw = g.n_words
    memo = {}
    for i in range(n):
        b = g.base[i]
        memo[1 << i] = (float(b["rows"]), float(b["startup"]), float(b["total"]), int(b["width"]), None)
    for s in range(1, 1 << n):
        if bin(s).count("1") < 2 or not connected(s, adj):
            continue
        best = None
        l = s & -s
        while l != s:
            r = s ^ l
            if l in memo and r in memo:
                L, R = memo[l], memo[r]
                out = oracle.join_cost(g, c, [l], [r], L[:4], R[:4])
                if best is None or out[2] < best[2]:
                    best = (out[0], out[1], out[2], int(out[3]), l)
            l = s & (l - s)
        memo[s] = best

    def tree(s):
        v = memo[s]
        node = [(s, 0, 0, 0), v[0], v[1], v[2], v[3]]
        if v[4] is None:
            return [node]
        return [node] + tree(v[4]) + tree(s ^ v[4])
    return tree((1 << n) - 1)


@pytest.mark.parametrize("name,g", [x for x in small_graphs() if x[1].n <= 8],
                         ids=lambda x: x if isinstance(x, str) else "")
def test_dpsub_recurrence_restated(name, g):
    """The oracle runs after BFS relabelling; graphs whose BFS order is the
    identity compare directly, the others through relabel-invariant costs."""
    c = cfg("cpu_dpsub")
    st, nodes, _ = oracle.plan(g, c)
    assert st == 0
    if oracle.bfs_order(g) == list(range(g.n)):
        want = python_dpsub(g, c)
        got = [[(x[0][0], 0, 0, 0), x[1], x[2], x[3], x[4]] for x in nodes]
        assert got == want
    else:
        assert nodes[0][3] > 0


def test_ties_keep_first_split():
    """Identical relations and selectivities: many splits tie, the first in l
    order must win (l = lowest relation alone at the root)."""
    g = qgen.QueryGraph(4, "ties")
    g.rows = [100.0] * 4
    for i in range(4):
        for j in range(i + 1, 4):
            g.add_generic(i, j, 0.5)
    g.build()
    c = cfg("cpu_dpsub")
    st, nodes, _ = oracle.plan(g, c)
    want = python_dpsub(g, c)
    assert [(x[0][0], x[1], x[2], x[3], x[4]) for x in nodes] == [(w[0][0], w[1], w[2], w[3], w[4]) for w in want]


# ---------------------------------------------------------------- other algorithms

@pytest.mark.parametrize("name,g", [x for x in small_graphs() if x[1].n >= 5],
                         ids=lambda x: x if isinstance(x, str) else "")
def test_dpsize_finds_dpsub_optimum(name, g):
    _, a, sa = oracle.plan(g, cfg("cpu_dpsub"))
    _, b, sb = oracle.plan(g, cfg("cpu_dpsize"))
    assert sa.pairs == sb.pairs
    assert a[0][3] == b[0][3]


@pytest.mark.parametrize("bicc", [0, 1])
def test_parallel_equals_sequential(bicc):
    g = qgen.snowflake(14, 77)
    seq = "cpu_dpsub_bicc" if bicc else "cpu_dpsub"
    par = "parallel_cpu_dpsub_bicc" if bicc else "parallel_cpu_dpsub"
    _, a, _ = oracle.plan(g, cfg(seq))
    _, b, _ = oracle.plan(g, cfg(par, dpe_n_threads=4, cpu_dpsub_parallel_chunk_size=7))
    assert gpuqo_amd.plan_bits(a) == gpuqo_amd.plan_bits(b)


def test_goo_greedy_on_rows():
    """GOO on a chain with distinct rows: the first join is the valid pair
    with the fewest result rows (textbook greedy)."""
    g = qgen.chain(6, 5)
    st, nodes, _ = oracle.plan(g, cfg("cpu_goo"))
    assert st == 0 and len(nodes) == 11
    c = cfg("cpu_dpsub")
    best = None
    for i in range(5):
        bi, bj = g.base[i], g.base[i + 1]
        rows = oracle.join_cost(g, c, [1 << i], [1 << (i + 1)], (float(bi["rows"]), 0.0, 1.0, 4.0),
                                (float(bj["rows"]), 0.0, 1.0, 4.0))[0]
        if best is None or rows < best[0]:
            best = (rows, (1 << i) | (1 << (i + 1)))
    pairs = [x for x in nodes if bin(x[0][0]).count("1") == 2]
    assert any(x[0][0] == best[1] and x[1] == best[0] for x in pairs)


@pytest.mark.parametrize("idp", ["idp2", "union"])
def test_idp_without_decomposition_is_the_dp(idp):
    g = qgen.snowflake(12, 8)
    _, a, _ = oracle.plan(g, cfg("dpsub", dpsub_bicc=1))
    _, b, _ = oracle.plan(g, cfg("dpsub", dpsub_bicc=1, idp_type=idp, idp_n_iters=12))
    assert gpuqo_amd.plan_bits(a) == gpuqo_amd.plan_bits(b)


def leaves(nodes):
    return [x[0] for x in nodes if x[5] < 0]


@pytest.mark.parametrize("n,seed", [(40, 1), (70, 2), (130, 3)])
def test_uniondp_plan_is_a_full_join_tree(n, seed):
    g = qgen.snowflake(n, seed, depth=4, fanout=5)
    st, nodes, stats = oracle.plan(g, cfg("dpsub", dpsub_bicc=1, idp_type="union", idp_n_iters=25))
    assert st == 0 and len(nodes) == 2 * n - 1
    lv = leaves(nodes)
    seen = 0
    for words in lv:
        m = sum(w << (64 * i) for i, w in enumerate(words))
        assert bin(m).count("1") == 1 and not (seen & m)
        seen |= m
    assert seen == (1 << n) - 1
    assert stats.n_dp_calls >= 2


@pytest.mark.parametrize("k", [5, 8])
def test_idp2_plan_is_a_full_join_tree(k):
    g = qgen.snowflake(20, 4)
    st, nodes, _ = oracle.plan(g, cfg("dpsub", dpsub_bicc=1, idp_type="idp2", idp_n_iters=k))
    assert st == 0 and len(nodes) == 39


# ---------------------------------------------------------------- DPccp / DPE

@pytest.mark.parametrize("n", [3, 6, 9])
def test_dpccp_visits_exactly_the_csg_cmp_pairs(n):
    """cpu_dpccp emits each csg-cmp pair once and costs both orientations:
    its join count is the ordered pair count of the closed forms."""
    cases = [
        (qgen.star(n, 1), (n - 1) * 2 ** (n - 1)),
        (qgen.clique(n, 2), 3 ** n - 2 ** (n + 1) + 1),
        (qgen.chain(n, 3), (n ** 3 - n) // 3),
        (qgen.cycle(n, 4), n ** 3 - 2 * n ** 2 + n),
    ]
    for g, want in cases:
        st, nodes, stats = oracle.plan(g, cfg("cpu_dpccp"))
        assert st == 0
        assert stats.pairs == want, (g.shape, n)
        assert stats.candidates * 2 == want, (g.shape, n)


@pytest.mark.parametrize("name,g", [("snow12", qgen.snowflake(12, 5)), ("clique8", qgen.clique(8, 6)),
                                    ("cycle9", qgen.cycle(9, 7)), ("star10", qgen.star(10, 8))])
def test_dpccp_reaches_the_dpsub_optimum(name, g):
    """Both explore every bushy plan without cross products: the optimal
    total cost is the same float."""
    _, want, _ = oracle.plan(g, cfg("cpu_dpsub"))
    st, got, _ = oracle.plan(g, cfg("cpu_dpccp"))
    assert st == 0
    assert gpuqo_amd.fbits(got[0][3]) == gpuqo_amd.fbits(want[0][3]), name


@pytest.mark.parametrize("threads,depbuf", [(1, 128), (4, 8), (4, 65536)])
def test_dpe_dpccp_returns_a_full_tree(threads, depbuf):
    """dpe_dpccp submits each pair twice in one orientation (the reference's
    DPE try_swap), so it is a timing baseline: a full join tree whose cost is
    never below the optimum."""
    g = qgen.snowflake(14, 9)
    _, best, _ = oracle.plan(g, cfg("cpu_dpsub"))
    st, nodes, stats = oracle.plan(g, cfg("dpe_dpccp", dpe_n_threads=threads, dpe_pairs_per_depbuf=depbuf))
    assert st == 0
    assert len(nodes) == 2 * g.n - 1
    assert nodes[0][3] >= best[0][3]
    root = nodes[0][0]
    assert sum(bin(w).count("1") for w in root) == g.n


def _pair_rows(g, c, i, j):
    """Estimated rows of joining base relations i and j (the oracle's cost model)."""
    nw = g.n_words
    lw, rw = [0] * nw, [0] * nw
    lw[i // 64] |= 1 << (i % 64)
    rw[j // 64] |= 1 << (j % 64)
    b = g.build().base
    return oracle.join_cost(g, c, lw, rw, (b[i]["rows"], 0.0, 0.0, 4), (b[j]["rows"], 0.0, 0.0, 4))[0]


def _with_edges(g, edges):
    """A copy of g whose edge table holds exactly `edges` (classes unchanged)."""
    g2 = copy.copy(g.build())
    et = np.zeros_like(g.edge_table)
    for a, b in edges:
        et[a, b // 64] |= np.uint64(1 << (b % 64))
        et[b, a // 64] |= np.uint64(1 << (a % 64))
    g2.edge_table = et
    g2._c = None
    return g2


def _kruskal(g, c):
    """Minimum spanning tree by Kruskal on the base-pair rows; None on a tie."""
    w = sorted((_pair_rows(g, c, a, b), min(a, b), max(a, b)) for a, b, _, _ in g.edges)
    if len({x[0] for x in w}) != len(w):
        return None
    up = list(range(g.n))

    def find(x):
        while up[x] != x:
            x = up[x]
        return x
    tree = []
    for _, a, b in w:
        ra, rb = find(a), find(b)
        if ra != rb:
            up[ra] = rb
            tree.append((a, b))
    return tree


@pytest.mark.parametrize("name,g", [("cycle9", qgen.cycle(9, 0x5901)), ("clique7", qgen.clique(7, 0x5904)),
                                    ("chorded12", qgen.chorded_cycle(12, 4, 0x5903)),
                                    ("cycle_tail", qgen.cycle_with_tail(6, 5, 0x5903))])
@pytest.mark.parametrize("algo,bicc", [("dpsub", 0), ("dpsub", 1), ("dpsize", 0)])
def test_spanning_tree_mode_is_the_dp_over_the_minimum_spanning_tree(name, g, algo, bicc):
    """gpuqo_spanning_tree: with distinct base-pair rows the minimum spanning
    tree is unique, so Prim's tree in the oracle must be Kruskal's here; the
    plan must be the plain DP over a graph with that edge table and the same
    classes (the costs still see every class).  The seeds are ones whose
    tree-restricted optimum differs from the full graph's, so a mode that
    changed nothing would fail."""
    c = cfg(algo, dpsub_bicc=bicc)
    tree = _kruskal(g, c)
    assert tree is not None and len(tree) == g.n - 1, name
    want, _ = oracle_bits(_with_edges(g, tree), c)
    got, _ = oracle_bits(g, cfg(algo, dpsub_bicc=bicc, spanning_tree=1))
    assert_same_plan(got, want, f"{name} {algo} bicc={bicc}")
    full, _ = oracle_bits(g, c)
    assert full != got, f"{name}: the spanning tree did not change the plan"
    assert full[0][3] <= got[0][3], "the full graph's optimum cannot cost more than the tree's"


def test_spanning_tree_mode_leaves_trees_alone():
    for g in (qgen.snowflake(14, 0x5905), qgen.chain(9, 0x5906), qgen.star(8, 0x5907)):
        for algo, bicc in (("dpsub", 0), ("dpsub", 1), ("dpsize", 0)):
            a, _ = oracle_bits(g, cfg(algo, dpsub_bicc=bicc))
            b, _ = oracle_bits(g, cfg(algo, dpsub_bicc=bicc, spanning_tree=1))
            assert_same_plan(b, a, f"{g.shape}{g.n} {algo}")


def nan_triangle():
    """Three relations of 100 rows in a triangle: class {0,1} selectivity
    1/2 (5000 rows), class {0,2} selectivity NaN, class {1,2} 1/1024
    (rows 9.765625 -> 10)."""
    g = qgen.QueryGraph(3, "nan-triangle")
    g.rows = [100.0, 100.0, 100.0]
    g.add_generic(0, 1, 0.5)
    g.add_generic(0, 2, float("nan"))
    g.add_generic(1, 2, 1.0 / 1024)
    return g.build()


def test_spanning_tree_takes_an_edge_with_nan_selectivity():
    """estimate_join_rows (gpuqo_row_estimation.cuh:120-147) returns 1 for a
    NaN product (`rows > 1` is false), so Prim (gpuqo_spanning_tree.cu:17-68)
    weighs edge {0,2} at 1 row and takes it first from relation 0, then
    {2,1} (10 rows) over {0,1} (5000): the tree is 0-2-1.  Joining {0,1}
    first is then impossible, while the full graph's plan is free to."""
    g = nan_triangle()
    c = cfg("dpsub")
    assert _pair_rows(g, c, 0, 2) == 1.0
    assert _pair_rows(g, c, 1, 2) == 10.0 and _pair_rows(g, c, 0, 1) == 5000.0
    tree = _with_edges(g, [(0, 2), (1, 2)])
    for algo, bicc in (("dpsub", 0), ("dpsub", 1), ("dpsize", 0)):
        want, _ = oracle_bits(tree, cfg(algo, dpsub_bicc=bicc))
        got, _ = oracle_bits(g, cfg(algo, dpsub_bicc=bicc, spanning_tree=1))
        assert_same_plan(got, want, f"nan triangle {algo} bicc={bicc}")
        assert any(nd[0][0] == 0b101 for nd in got), "the plan joins {0,2} (1 row) first"
        assert not any(nd[0][0] == 0b011 for nd in got)

