"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  The product never does.
"""
import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        from gpuqo_amd import abi
        L = C.CDLL(_LIB)
        L.oracle_plan.argtypes = [C.POINTER(abi.Query), C.POINTER(abi.Config), C.POINTER(abi.PlanOut)]
        L.oracle_plan.restype = C.c_int
        L.oracle_plan_connected.argtypes = [C.POINTER(abi.Query), C.POINTER(abi.Config), C.POINTER(abi.PlanOut)]
        L.oracle_plan_connected.restype = C.c_int
        L.oracle_join_cost.argtypes = [C.POINTER(abi.Query), C.POINTER(abi.Config), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]
        L.oracle_join_cost.restype = C.c_int
        L.oracle_hash_table_size.argtypes = [C.c_float, C.c_int, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.oracle_bfs_order.argtypes = [C.POINTER(abi.Query), C.POINTER(C.c_int)]
        L.oracle_bfs_order.restype = C.c_int
        L.oracle_blocks.argtypes = [C.POINTER(abi.Query), C.c_uint64, C.POINTER(C.c_uint64), C.c_int]
        L.oracle_blocks.restype = C.c_int
        L.oracle_dpsize_level1_order.argtypes = [C.c_int, C.POINTER(C.c_int)]
        L.oracle_dpsize_level1_order.restype = C.c_int
        L.oracle_cost_probe.argtypes = [C.POINTER(abi.Query), C.POINTER(abi.Config), C.c_int,
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_float),
                                        C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint64,
                                        C.POINTER(C.c_float)]
        L.oracle_cost_probe.restype = C.c_int
        L.oracle_hash_bucketsize.argtypes = [C.POINTER(abi.VarInfo), C.c_float, C.c_float, C.c_int]
        L.oracle_hash_bucketsize.restype = C.c_float
        L.oracle_dp_first_levels.argtypes = [C.POINTER(abi.Query), C.POINTER(abi.Config), C.c_int,
                                             C.POINTER(abi.PlanStats)]
        L.oracle_dp_first_levels.restype = C.c_int
        _lib = L
    return _lib


def plan(query, cfg):
    """Run the oracle; returns (status, nodes, stats) with nodes as tuples
    (relids, rows, startup, total, width, left, right)."""
    from gpuqo_amd import abi, plan_tuples
    n = query.n
    cap = 2 * n
    nodes = (abi.PlanNode * cap)()
    out = abi.PlanOut(C.cast(nodes, C.POINTER(abi.PlanNode)), cap, 0)
    st = lib().oracle_plan(C.byref(query.to_c()), C.byref(cfg), C.byref(out))
    return st, plan_tuples(out), out.stats


def plan_connected(query, cfg):
    """plan() with DPsub / MPDP visiting only the connected sets, size by
    size (oracle_plan_connected): the same plans, for DPs above 32
    relations."""
    from gpuqo_amd import abi, plan_tuples
    n = query.n
    cap = 2 * n
    nodes = (abi.PlanNode * cap)()
    out = abi.PlanOut(C.cast(nodes, C.POINTER(abi.PlanNode)), cap, 0)
    st = lib().oracle_plan_connected(C.byref(query.to_c()), C.byref(cfg), C.byref(out))
    return st, plan_tuples(out), out.stats


def join_cost(query, cfg, l_words, r_words, lvals, rvals):
    nw = query.n_words
    lw = (C.c_uint64 * nw)(*l_words)
    rw = (C.c_uint64 * nw)(*r_words)
    lv = (C.c_float * 4)(*lvals)
    rv = (C.c_float * 4)(*rvals)
    ov = (C.c_float * 4)()
    st = lib().oracle_join_cost(C.byref(query.to_c()), C.byref(cfg), lw, rw, lv, rv, ov)
    assert st == 0
    return tuple(ov)


def dp_first_levels(query, cfg, max_levels):
    """Run the configured DP through levels 2..max_levels only (a bounded
    timing sample); returns the stats (pairs, sets, wall_ms)."""
    from gpuqo_amd import abi
    st = abi.PlanStats()
    rc = lib().oracle_dp_first_levels(C.byref(query.to_c()), C.byref(cfg), max_levels, C.byref(st))
    assert rc == 0
    return st


PROBE_JOIN, PROBE_NESTLOOP, PROBE_HASHJOIN, PROBE_BUCKET_STATS, PROBE_INNER_UNIQUE, PROBE_BASEREL = range(6)


def cost_probe(query, cfg, which, l_words, r_words, lvals=(0, 0, 0, 0), rvals=(0, 0, 0, 0), extra=(0, 0, 0, 0, 0),
               composite_mask=0):
    """One branch of the oracle's cost model (see oracle_cost_probe):
    which = PROBE_JOIN (extra[2] = join rows), PROBE_NESTLOOP / PROBE_HASHJOIN
    (extra = inner_unique, indexed_join_quals, joinrows, outer_match_frac,
    match_count), PROBE_BUCKET_STATS (extra[0] = nbuckets),
    PROBE_INNER_UNIQUE or PROBE_BASEREL (cost_baserel of l's lowest
    relation).  Returns two floats (one for PROBE_INNER_UNIQUE)."""
    nw = query.n_words
    lw = (C.c_uint64 * nw)(*l_words)
    rw = (C.c_uint64 * nw)(*r_words)
    lv = (C.c_float * 4)(*lvals)
    rv = (C.c_float * 4)(*rvals)
    ex = (C.c_float * 5)(*extra)
    out = (C.c_float * 2)()
    st = lib().oracle_cost_probe(C.byref(query.to_c()), C.byref(cfg), which, lw, rw, lv, rv, ex, composite_mask, out)
    assert st == 0
    return out[0] if which == PROBE_INNER_UNIQUE else (out[0], out[1])


def hash_bucketsize(stadistinct, stanullfrac, mcvfreq, rows, tuples, nbuckets):
    from gpuqo_amd import abi
    v = abi.VarInfo()
    v.stadistinct, v.stanullfrac, v.mcvfreq = stadistinct, stanullfrac, mcvfreq
    return lib().oracle_hash_bucketsize(C.byref(v), rows, tuples, nbuckets)


def hash_table_size(ntuples, width, work_mem=4096):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    lib().oracle_hash_table_size(ntuples, width, work_mem, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def bfs_order(query):
    out = (C.c_int * query.n)()
    k = lib().oracle_bfs_order(C.byref(query.to_c()), out)
    return list(out)[:k]


def blocks(query, set0):
    out = (C.c_uint64 * 64)()
    k = lib().oracle_blocks(C.byref(query.to_c()), set0, out, 64)
    return list(out)[:k]


def dpsize_level1_order(n):
    out = (C.c_int * n)()
    k = lib().oracle_dpsize_level1_order(n, out)
    return list(out)[:k]
