"""Shared test helpers: run the GPU path and the oracle, compare bit-exactly."""
import gpuqo_amd
import oracle
from gpuqo_amd import abi, qgen


def cfg(algo="dpsub", bicc=0, **kw):
    return abi.default_config(algorithm=algo, dpsub_bicc=bicc, **kw)


def oracle_bits(g, c):
    st, nodes, stats = oracle.plan(g, c)
    assert st == 0, f"oracle status {st}"
    return gpuqo_amd.plan_bits(nodes), stats


def gpu_bits(g, c):
    nodes, stats = gpuqo_amd.plan(g, c)
    return gpuqo_amd.plan_bits(nodes), stats


def assert_same_plan(a, b, what=""):
    assert len(a) == len(b), f"{what}: {len(a)} vs {len(b)} nodes"
    for i, (x, y) in enumerate(zip(a, b)):
        assert x == y, f"{what}: node {i} differs: {x} vs {y}"


def small_graphs():
    """Seeded small join graphs of every shape (name, graph)."""
    out = []
    for n in (2, 3, 5, 8):
        out.append((f"star{n}", qgen.star(n, 11 + n)))
        out.append((f"chain{n}", qgen.chain(n, 21 + n)))
        out.append((f"cycle{n}", qgen.cycle(n, 31 + n)))
        out.append((f"clique{n}", qgen.clique(n, 41 + n)))
        out.append((f"snow{n}", qgen.snowflake(n, 51 + n)))
    return out


def _figure8(n, seed):
    """Two cycles sharing relation 0."""
    a = n // 2
    edges = [(i, i + 1) for i in range(a - 1)] + [(a - 1, 0)]
    prev = 0
    for i in range(a, n):
        edges.append((prev, i))
        prev = i
    edges.append((n - 1, 0))
    return qgen.graph(n, edges, seed, "figure8")


def _triangle_ring(t, seed):
    """t triangles in a ring, consecutive ones sharing a relation: one
    biconnected block of 2t relations."""
    n = 2 * t
    edges = []
    for i in range(t):
        a, b, c = 2 * i, 2 * i + 1, (2 * i + 2) % n
        edges += [(a, b), (b, c), (a, c)]
    return qgen.graph(n, edges, seed, "triangle-ring")


def dpsub_sparse_graphs():
    """Sparse cyclic graphs of 16-22 relations (name, graph) whose DPsub splits
    are found through their biconnected blocks, and the triangle ring, one
    16-relation block, which keeps the subset scan."""
    return [("cycle18", qgen.cycle(18, 0xb18)), ("cycle22", qgen.cycle(22, 0xb22)),
            ("cycle12+tail6", qgen.cycle_with_tail(12, 6, 0xb21)), ("figure8-17", _figure8(17, 0xb24)),
            ("triangle-ring16", _triangle_ring(8, 0xb23))]


def tree_pairs_and_sets(g):
    """Ordered csg-cmp pairs and connected sets of two or more relations of
    a tree query, from the subtree polynomials P_v(x) = x * prod over the
    children c of (1 + P_c(x)): every connected set of k relations has
    2(k - 1) ordered splits."""
    adj = g.adjacency()
    order, parent = [0], {0: -1}
    for v in order:
        for w in range(g.n):
            if (adj[v] >> w) & 1 and w not in parent:
                parent[w] = v
                order.append(w)
    poly, total = {}, {}
    for v in reversed(order):
        p = {1: 1}
        for c in (w for w in order if parent[w] == v):
            q = dict(poly[c])
            q[0] = q.get(0, 0) + 1
            r = {}
            for a, x in p.items():
                for b, y in q.items():
                    r[a + b] = r.get(a + b, 0) + x * y
            p = r
        poly[v] = p
        for k, m in p.items():
            total[k] = total.get(k, 0) + m
    return (sum(2 * (k - 1) * m for k, m in total.items() if k >= 2),
            sum(m for k, m in total.items() if k >= 2))
