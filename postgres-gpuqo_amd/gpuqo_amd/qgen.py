"""Seeded synthetic join-graph generator ("qgen v1").

Shapes follow the reference's synthetic schemas (scripts/databases/
{star,snowflake,clique,chain}-small/*.py) but are emitted directly as
planner input instead of SQL.  Every value is a deterministic function of
the seed:

* RNG: xoshiro256** seeded by four SplitMix64 outputs of the seed;
  u = (next() >> 11) * 2**-53.
* logU[a, b] = exp(ln a + u (ln b - ln a)) in double; rows = float(rint(x));
  tuples = rows.
* width = 4 * (1 + number of FK columns the relation holds).
* base cost = {0, float(pages + 0.01 * tuples)},
  pages = ceil(tuples / floor(8168 / (28 + MAXALIGN(width)))).
* one 2-member equivalence class per edge, in edge-creation order.
* PK-FK edge (holder -> pk): fks[holder] = {pk}, sel = float(1 / rows_pk),
  vars[pk] = unique index (rows 1, cost {0.29, 8.31}),
  vars[holder].stadistinct = min(rows_pk, rows_holder).
* generic edge with selectivity s: both stadistinct = min(rows, 1/s).
"""
import math

import numpy as np

from . import abi

M64 = (1 << 64) - 1


class Rng:
    """xoshiro256** seeded with SplitMix64."""

    def __init__(self, seed):
        x = seed & M64
        s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & M64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            s.append(z ^ (z >> 31))
        self.s = s

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & M64

    def next(self):
        s = self.s
        r = (self._rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return r

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def uint(self, a, b):
        """integer uniform on {a..b}"""
        return a + int(self.uniform() * (b - a + 1))

    def real(self, a, b):
        return a + self.uniform() * (b - a)

    def logu(self, a, b):
        la, lb = math.log(a), math.log(b)
        return math.exp(la + self.uniform() * (lb - la))


def f32(x):
    return float(np.float32(x))


def rows_of(x):
    return f32(np.rint(x))


class QueryGraph:
    """A join graph in the layout of gq_query (relation i = bit i)."""

    def __init__(self, n, shape=""):
        self.n = n
        self.shape = shape
        self.rows = [0.0] * n
        self.edges = []          # (holder_or_u, pk_or_v, kind, sel)
        self.fk_count = [0] * n

    def add_pkfk(self, holder, pk):
        sel = f32(1.0 / self.rows[pk])
        self.edges.append((holder, pk, "fk", sel))
        self.fk_count[holder] += 1

    def add_generic(self, u, v, sel):
        self.edges.append((u, v, "gen", f32(sel)))

    def build(self):
        """Freeze into numpy arrays (cached); returns self."""
        n = self.n
        nw = (n + 63) // 64
        base = np.zeros(n, dtype=[("rows", "<f4"), ("tuples", "<f4"), ("startup", "<f4"),
                                  ("total", "<f4"), ("width", "<i4")])
        for i in range(n):
            rows = self.rows[i]
            width = 4 * (1 + self.fk_count[i])
            maxalign = (width + 7) & ~7
            per_page = math.floor(8168 / (28 + maxalign))
            pages = math.ceil(rows / per_page)
            base[i] = (rows, rows, 0.0, f32(pages + 0.01 * rows), width)
        et = np.zeros((n, nw), dtype=np.uint64)
        offs = [0]
        members, sels, fks, vars_ = [], [], [], []
        for (a, b, kind, sel) in self.edges:
            et[a, b // 64] |= np.uint64(1 << (b % 64))
            et[b, a // 64] |= np.uint64(1 << (a % 64))
            lo, hi = min(a, b), max(a, b)
            members += [lo, hi]
            offs.append(len(members))
            sels.append(sel)
            for m in (lo, hi):
                fk = [0] * nw
                var = dict(stadistinct=0.0, stanullfrac=0.0, mcvfreq=0.0, index_available=0,
                           index_unique=0, index_rows=0.0, index_startup=0.0, index_total=0.0)
                if kind == "fk":
                    holder, pk = a, b
                    if m == holder:
                        fk[pk // 64] |= 1 << (pk % 64)
                        var["stadistinct"] = min(self.rows[pk], self.rows[holder])
                    else:
                        var.update(stadistinct=-1.0, index_available=1, index_unique=1,
                                   index_rows=1.0, index_startup=f32(0.29), index_total=f32(8.31))
                else:
                    var["stadistinct"] = f32(min(self.rows[m], 1.0 / float(sel)))
                fks.append(fk)
                vars_.append(var)
        vdt = np.dtype([("stadistinct", "<f4"), ("stanullfrac", "<f4"), ("mcvfreq", "<f4"),
                        ("index_available", "u1"), ("index_unique", "u1"), ("pad_", "u1", (2,)),
                        ("index_rows", "<f4"), ("index_startup", "<f4"), ("index_total", "<f4")])
        va = np.zeros(len(vars_), dtype=vdt)
        for i, v in enumerate(vars_):
            for k, x in v.items():
                va[i][k] = x
        self.base = base
        self.edge_table = et
        self.n_words = nw
        self.ec_offsets = np.array(offs, dtype=np.int32)
        self.ec_members = np.array(members if members else [0], dtype=np.int32)
        self.ec_sels = np.array(sels if sels else [0.0], dtype=np.float32)
        self.ec_fks = np.array(fks if fks else [[0] * nw], dtype=np.uint64).reshape(-1, nw)
        self.ec_vars = va if len(va) else np.zeros(1, dtype=vdt)
        self.n_ec = len(self.edges)
        self._c = None
        return self

    def to_c(self):
        """gq_query view over the numpy arrays (kept alive by self)."""
        if self._c is None:
            if not hasattr(self, "base"):
                self.build()
            q = abi.Query()
            q.n_rels = self.n
            q.n_words = self.n_words
            q.base_rels = self.base.ctypes.data_as(abi.C.POINTER(abi.BaseRel))
            q.edge_table = self.edge_table.ctypes.data_as(abi.C.POINTER(abi.C.c_uint64))
            q.n_eq_classes = self.n_ec
            q.ec_member_offsets = self.ec_offsets.ctypes.data_as(abi.C.POINTER(abi.C.c_int32))
            q.ec_members = self.ec_members.ctypes.data_as(abi.C.POINTER(abi.C.c_int32))
            q.ec_sels = self.ec_sels.ctypes.data_as(abi.C.POINTER(abi.C.c_float))
            q.ec_fks = self.ec_fks.ctypes.data_as(abi.C.POINTER(abi.C.c_uint64))
            q.ec_vars = self.ec_vars.ctypes.data_as(abi.C.POINTER(abi.VarInfo))
            self._c = q
        return self._c

    def adjacency(self):
        adj = [0] * self.n
        for (a, b, _, _) in self.edges:
            adj[a] |= 1 << b
            adj[b] |= 1 << a
        return adj


def star(n, seed):
    """One fact logU[1e6,1e8] holding FKs to n-1 dimensions logU[1e2,1e6]."""
    r = Rng(seed)
    g = QueryGraph(n, "star")
    g.rows[0] = rows_of(r.logu(1e6, 1e8))
    for i in range(1, n):
        g.rows[i] = rows_of(r.logu(1e2, 1e6))
    for i in range(1, n):
        g.add_pkfk(0, i)
    return g.build()


def snowflake(n, seed, depth=3, fanout=4, root=(1e7, 1e9)):
    """Rooted tree of depth <= `depth`, fanout U{1..fanout}, BFS order,
    truncated at n nodes; child rows = max(10, parent / U[10, 1000]); the
    parent holds the FK.  Too few nodes: redraw with seed + 1."""
    s = seed
    while True:
        r = Rng(s)
        g = QueryGraph(n, "snowflake")
        g.rows[0] = rows_of(r.logu(*root))
        queue = [(0, 0)]
        cnt = 1
        qi = 0
        while qi < len(queue) and cnt < n:
            p, d = queue[qi]
            qi += 1
            if d >= depth:
                continue
            f = r.uint(1, fanout)
            for _ in range(f):
                if cnt >= n:
                    break
                c = cnt
                cnt += 1
                g.rows[c] = rows_of(max(10.0, g.rows[p] / r.real(10.0, 1000.0)))
                g.add_pkfk(p, c)
                queue.append((c, d + 1))
        if cnt == n:
            return g.build()
        s += 1


def clique(n, seed):
    """Complete graph, rows logU[1e3,1e7], selectivity logU[1e-6,1e-1]."""
    r = Rng(seed)
    g = QueryGraph(n, "clique")
    for i in range(n):
        g.rows[i] = rows_of(r.logu(1e3, 1e7))
    for i in range(n):
        for j in range(i + 1, n):
            g.add_generic(i, j, r.logu(1e-6, 1e-1))
    return g.build()


def chain(n, seed, cycle=False):
    """Path (or cycle), rows logU[1e3,1e7], selectivity logU[1e-5,1e-1]."""
    r = Rng(seed)
    g = QueryGraph(n, "cycle" if cycle else "chain")
    for i in range(n):
        g.rows[i] = rows_of(r.logu(1e3, 1e7))
    for i in range(n - 1):
        g.add_generic(i, i + 1, r.logu(1e-5, 1e-1))
    if cycle and n > 2:
        g.add_generic(n - 1, 0, r.logu(1e-5, 1e-1))
    return g.build()


def cycle(n, seed):
    return chain(n, seed, cycle=True)


def graph(n, edges, seed, shape="graph"):
    """Arbitrary join graph over the given (u, v) edges: rows logU[1e3,1e7],
    selectivity logU[1e-5,1e-1] per edge, in edge order."""
    r = Rng(seed)
    g = QueryGraph(n, shape)
    for i in range(n):
        g.rows[i] = rows_of(r.logu(1e3, 1e7))
    for (u, v) in edges:
        g.add_generic(u, v, r.logu(1e-5, 1e-1))
    return g.build()


def hub_chains(n, arms, seed):
    """A hub (relation 0) with `arms` chains of relations hanging off it, the
    n - 1 others dealt out evenly (the first arms one longer), built chain
    by chain; rows logU[1e3,1e7], selectivity logU[1e-5,1e-1] per edge."""
    edges = []
    nxt = 1
    per, extra = divmod(n - 1, arms)
    for a in range(arms):
        prev = 0
        for _ in range(per + (1 if a < extra else 0)):
            edges.append((prev, nxt))
            prev = nxt
            nxt += 1
    return graph(n, edges, seed, "hub-chains")


def cycle_with_tail(c, t, seed):
    """A cycle of c relations with a path of t more hanging off relation 0."""
    edges = [(i, i + 1) for i in range(c - 1)] + [(c - 1, 0)]
    prev = 0
    for i in range(c, c + t):
        edges.append((prev, i))
        prev = i
    return graph(c + t, edges, seed, "cycle+tail")


def chorded_cycle(n, step, seed):
    """A cycle with chords (i, i + step): one biconnected block."""
    edges = [(i, (i + 1) % n) for i in range(n)]
    edges += [(i, i + step) for i in range(0, n - step, 2)]
    return graph(n, edges, seed, "chorded-cycle")


SEED_BASE = 0x6770713000000000


WORKLOAD_SEEDS = {"star12": 1, "snowflake20": 2, "clique18": 3, "batch256": 4, "serve256": 5,
                  "tune": 6}           # tune + i, i = 0..5: the autotuner's sample queries


def workload_seed(name):
    """Seed of a named benchmark workload (fixed: the golden fixtures and the
    published numbers are tied to the graphs it generates)."""
    return SEED_BASE + WORKLOAD_SEEDS[name]


def query_star12():
    """12-relation star, plain DPsub."""
    return star(12, workload_seed("star12"))


def query_snowflake20():
    """20-relation snowflake (depth 3, fanout 1..4), MPDP."""
    return snowflake(20, workload_seed("snowflake20"))


def query_clique18():
    """18-relation clique, plain DPsub and DPsize."""
    return clique(18, workload_seed("clique18"))


def mixed_batch(count, seed, nmin=8, nmax=22):
    """Mixed shapes: 40% snowflake, 30% star, 15% chain, 15% cycle; query i
    draws its shape and size from Rng(seed + i)."""
    out = []
    for i in range(count):
        r = Rng(seed + i)
        n = r.uint(nmin, nmax)
        u = r.uniform()
        qs = (seed + i) ^ 0x5A5A5A5A
        if u < 0.40:
            out.append(snowflake(n, qs))
        elif u < 0.70:
            out.append(star(n, qs))
        elif u < 0.85:
            out.append(chain(n, qs))
        else:
            out.append(cycle(n, qs))
    return out


def queries_batch256(count=256):
    """The 256-query mixed batch (n in 8..22), MPDP."""
    return mixed_batch(count, workload_seed("batch256"))


def serving_mix(count, seed):
    """70%: snowflake 10-30 rels (direct MPDP); 30%: snowflake 60-200 rels,
    depth 4, fanout 1..5 (UnionDP k=25).  Returns (query, is_large)."""
    out = []
    for i in range(count):
        r = Rng(seed + i)
        u = r.uniform()
        qs = (seed + i) ^ 0x3C3C3C3C
        if u < 0.70:
            out.append((snowflake(r.uint(10, 30), qs), False))
        else:
            out.append((snowflake(r.uint(60, 200), qs, depth=4, fanout=5), True))
    return out
