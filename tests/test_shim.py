"""The Postgres drop-in, postgres-gpuqo_amd/shim/gpuqo_run_amd.c.

The shim is compiled against test stand-ins of the Postgres headers it
includes (tests/shim/include: palloc, Bitmapset, elog, the GUC globals and
the planner-info structs of src/include/optimizer/gpuqo_planner_info.h:15-57
and gpuqo_common.h:38-113) and linked with libgpuqo_amd.so into
tests/shim/libpg_shim_harness.so.  The tests build a GpuqoPlannerInfoC the
way gpuqo_main.c:380-500 lays it out -- relation i as Bitmapset member i+1,
one EqClassInfo per class in a linked list, sels in eqClassIndex order, one
fk set and one VarInfo per member -- with Python code of their own, and
check:

  * CPU: the ctypes mirror of the structs has the C layout; the shim's
    conversion gives exactly the generator's gq_query arrays (star-12,
    snowflake-20, a clique, a 70-relation query with a 3-member class, NULL
    edge and fk sets, two words per set); malformed infos are refused; every
    GUC lands in its own gq_config field and nothing else; a plan becomes a
    palloc'd pre-order QueryTreeC with member-(i+1) ids and NULL leaf
    children; gpuqo_run returns NULL with a WARNING for a CPU algorithm.
  * GPU: gpuqo_run returns the tree of gq_plan for the same query and GUCs
    (star-12, snowflake-20, DPsub/MPDP/DPsize/UnionDP), and an interrupt
    makes it return NULL after CHECK_FOR_INTERRUPTS().
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import gpuqo_amd
import oracle
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM_DIR = os.path.join(HERE, "shim")
SHIM_LIB = os.path.join(SHIM_DIR, "libpg_shim_harness.so")


class PathCost(C.Structure):
    _fields_ = [("startup", C.c_float), ("total", C.c_float)]


class IndexInfo(C.Structure):
    _fields_ = [("available", C.c_bool), ("unique", C.c_bool), ("rows", C.c_float), ("cost", PathCost)]


class VarStats(C.Structure):
    _fields_ = [("stadistinct", C.c_float), ("stanullfrac", C.c_float), ("mcvfreq", C.c_float)]


class PgVarInfo(C.Structure):
    _fields_ = [("stats", VarStats), ("index", IndexInfo)]


class EqClassInfo(C.Structure):
    pass


EqClassInfo._fields_ = [("eclass", C.c_void_p), ("relids", C.c_void_p), ("sels", C.POINTER(C.c_float)),
                        ("vars", C.POINTER(PgVarInfo)), ("fk", C.POINTER(C.c_void_p)),
                        ("next", C.POINTER(EqClassInfo))]


class BaseRelationC(C.Structure):
    _fields_ = [("id", C.c_void_p), ("rows", C.c_float), ("tuples", C.c_float), ("cost", PathCost),
                ("width", C.c_int)]


class PlannerInfoC(C.Structure):
    _fields_ = [("n_rels", C.c_int), ("base_rels", C.POINTER(BaseRelationC)),
                ("edge_table", C.POINTER(C.c_void_p)), ("eq_classes", C.POINTER(EqClassInfo)),
                ("n_eq_classes", C.c_int), ("n_eq_class_sels", C.c_int), ("n_eq_class_fks", C.c_int),
                ("n_eq_class_vars", C.c_int)]


class QueryTreeC(C.Structure):
    pass


QueryTreeC._fields_ = [("id", C.c_void_p), ("rows", C.c_float), ("cost", PathCost), ("width", C.c_int),
                       ("left", C.POINTER(QueryTreeC)), ("right", C.POINTER(QueryTreeC))]

_shim = None


def shim():
    global _shim
    if _shim is None:
        if not os.path.exists(SHIM_LIB):
            subprocess.run(["make", "-s", "-C", SHIM_DIR], check=True)
        gpuqo_amd.lib()                       # the libgpuqo_amd.so the harness links
        L = C.CDLL(SHIM_LIB)
        L.pg_standin_bms_make.argtypes = [C.POINTER(C.c_uint64), C.c_int]
        L.pg_standin_bms_make.restype = C.c_void_p
        L.bms_next_member.argtypes = [C.c_void_p, C.c_int]
        L.bms_next_member.restype = C.c_int
        L.pg_standin_last_message.argtypes = [C.POINTER(C.c_int)]
        L.pg_standin_last_message.restype = C.c_char_p
        L.pg_standin_palloc_count.restype = C.c_long
        L.pg_standin_layout.argtypes = [C.POINTER(C.c_long), C.c_int]
        L.gpuqo_amd_query_from_info.argtypes = [C.POINTER(PlannerInfoC), C.POINTER(abi.Query)]
        L.gpuqo_amd_config_from_gucs.argtypes = [C.c_int, C.POINTER(abi.Config)]
        L.gpuqo_amd_config_from_gucs.restype = None
        L.gpuqo_amd_tree_from_plan.argtypes = [C.POINTER(abi.PlanOut)]
        L.gpuqo_amd_tree_from_plan.restype = C.POINTER(QueryTreeC)
        L.gpuqo_run.argtypes = [C.c_int, C.POINTER(PlannerInfoC)]
        L.gpuqo_run.restype = C.POINTER(QueryTreeC)
        _shim = L
    return _shim


@pytest.fixture(autouse=True)
def _fresh_context():
    """Each test starts from an empty palloc arena and a clear log, as a
    planner call starts in a fresh memory context."""
    L = shim()
    L.pg_standin_reset()
    L.pg_standin_clear_log()
    yield
    L.pg_standin_reset()


# ---------------------------------------------------------------- the input

class Spec:
    """A query in plain Python terms: relation i, relation sets as Python
    sets, classes as (members ascending, sels in eqClassIndex order, fk set
    per member, VarInfo dict per member)."""

    def __init__(self, n, base, edges, classes):
        self.n, self.base, self.edges, self.classes = n, base, edges, classes


def spec_of(g):
    """The Spec of a generator graph, read from its gq_query arrays."""
    n, nw = g.n, g.n_words

    def rels(words):
        return {i for i in range(64 * nw) if (int(words[i // 64]) >> (i % 64)) & 1}
    base = [tuple(float(g.base[i][k]) for k in ("rows", "tuples", "startup", "total")) + (int(g.base[i]["width"]),)
            for i in range(n)]
    edges = [rels(g.edge_table[i]) for i in range(n)]
    classes, so = [], 0
    for e in range(g.n_ec):
        a, b = int(g.ec_offsets[e]), int(g.ec_offsets[e + 1])
        s = b - a
        mem = [int(x) for x in g.ec_members[a:b]]
        sels = [float(x) for x in g.ec_sels[so:so + s * (s - 1) // 2]]
        so += s * (s - 1) // 2
        fks = [rels(g.ec_fks[m]) for m in range(a, b)]
        vs = [{k: g.ec_vars[m][k].item() for k in ("stadistinct", "stanullfrac", "mcvfreq", "index_available",
                                                    "index_unique", "index_rows", "index_startup", "index_total")}
              for m in range(a, b)]
        classes.append((mem, sels, fks, vs))
    return Spec(n, base, edges, classes)


def bms(rels, keep):
    """A Bitmapset with member r+1 for each relation r (NULL when empty)."""
    if not rels:
        return None
    top = max(rels) + 1
    words = (C.c_uint64 * (top // 64 + 1))()
    for r in rels:
        words[(r + 1) // 64] |= 1 << ((r + 1) % 64)
    keep.append(words)
    return shim().pg_standin_bms_make(words, len(words))


def planner_info(spec):
    """The GpuqoPlannerInfoC of `spec`; returns (info, keepalive)."""
    keep = []
    base = (BaseRelationC * spec.n)()
    et = (C.c_void_p * spec.n)()
    for i, (rows, tuples, st, tot, w) in enumerate(spec.base):
        base[i].id = bms({i}, keep)
        base[i].rows, base[i].tuples, base[i].width = rows, tuples, w
        base[i].cost.startup, base[i].cost.total = st, tot
        et[i] = bms(spec.edges[i], keep)
    nodes = [EqClassInfo() for _ in spec.classes]
    n_sels = n_vars = 0
    for ec, (mem, sels, fks, vs) in zip(nodes, spec.classes):
        s = len(mem)
        assert mem == sorted(mem) and len(sels) == s * (s - 1) // 2
        ec.relids = bms(set(mem), keep)
        ec.sels = (C.c_float * max(1, len(sels)))(*sels)
        fk = (C.c_void_p * s)(*[bms(f, keep) for f in fks])
        ec.fk = C.cast(fk, C.POINTER(C.c_void_p))
        va = (PgVarInfo * s)()
        for j, v in enumerate(vs):
            va[j].stats.stadistinct, va[j].stats.stanullfrac, va[j].stats.mcvfreq = (
                v["stadistinct"], v["stanullfrac"], v["mcvfreq"])
            va[j].index.available, va[j].index.unique = bool(v["index_available"]), bool(v["index_unique"])
            va[j].index.rows = v["index_rows"]
            va[j].index.cost.startup, va[j].index.cost.total = v["index_startup"], v["index_total"]
        ec.vars = C.cast(va, C.POINTER(PgVarInfo))
        keep += [ec.sels, fk, va]
        n_sels += len(sels)
        n_vars += s
    for a, b in zip(nodes, nodes[1:]):
        a.next = C.pointer(b)
    info = PlannerInfoC(spec.n, C.cast(base, C.POINTER(BaseRelationC)), C.cast(et, C.POINTER(C.c_void_p)),
                        C.pointer(nodes[0]) if nodes else None, len(nodes), n_sels, n_vars, n_vars)
    keep += [base, et, nodes]
    return info, keep


def converted(info):
    """The shim's gq_query of `info`, read back into numpy arrays."""
    q = abi.Query()
    assert shim().gpuqo_amd_query_from_info(C.byref(info), C.byref(q)) == 0
    n, nw, e = q.n_rels, q.n_words, q.n_eq_classes
    moff = np.ctypeslib.as_array(q.ec_member_offsets, (e + 1,)).copy()
    m = int(moff[-1])
    s_tot = sum((int(moff[i + 1] - moff[i]) * int(moff[i + 1] - moff[i] - 1)) // 2 for i in range(e))
    base = [(q.base_rels[i].rows, q.base_rels[i].tuples, q.base_rels[i].cost.startup, q.base_rels[i].cost.total,
             q.base_rels[i].width) for i in range(n)]
    vdt = qgen.QueryGraph(1).build().ec_vars.dtype
    return dict(n=n, nw=nw, n_ec=e, base=base,
                edge_table=np.ctypeslib.as_array(q.edge_table, (n, nw)).copy(),
                moff=moff, members=np.ctypeslib.as_array(q.ec_members, (max(1, m),))[:m].copy(),
                sels=np.ctypeslib.as_array(q.ec_sels, (max(1, s_tot),))[:s_tot].copy(),
                fks=np.ctypeslib.as_array(q.ec_fks, (max(1, m) * nw,)).reshape(-1, nw)[:m].copy(),
                vars=np.frombuffer(C.string_at(q.ec_vars, m * vdt.itemsize), dtype=vdt).copy())


# ------------------------------------------------------------------ CPU

def test_harness_structs_have_the_c_layout():
    out = (C.c_long * 64)()
    k = shim().pg_standin_layout(out, 64)
    want = [C.sizeof(PgVarInfo), PgVarInfo.index.offset,
            IndexInfo.unique.offset, IndexInfo.rows.offset, IndexInfo.cost.offset,
            C.sizeof(EqClassInfo), EqClassInfo.relids.offset, EqClassInfo.sels.offset,
            EqClassInfo.vars.offset, EqClassInfo.fk.offset, EqClassInfo.next.offset,
            C.sizeof(BaseRelationC), BaseRelationC.rows.offset, BaseRelationC.cost.offset,
            BaseRelationC.width.offset,
            C.sizeof(PlannerInfoC), PlannerInfoC.base_rels.offset, PlannerInfoC.edge_table.offset,
            PlannerInfoC.eq_classes.offset, PlannerInfoC.n_eq_classes.offset, PlannerInfoC.n_eq_class_vars.offset,
            C.sizeof(QueryTreeC), QueryTreeC.rows.offset, QueryTreeC.cost.offset, QueryTreeC.width.offset,
            QueryTreeC.left.offset, QueryTreeC.right.offset]
    assert list(out)[:k] == want


def test_one_libgpuqo_amd_in_the_process():
    """The harness's libgpuqo_amd.so (found through its rpath) is the one
    the binding loaded: one interrupt flag, one device context."""
    shim()
    ids = set()
    with open("/proc/self/maps") as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 6 and parts[5].endswith("/libgpuqo_amd.so"):
                ids.add((parts[3], parts[4]))
    assert len(ids) == 1, ids


CONVERT_CASES = [("star12", qgen.query_star12), ("snowflake20", qgen.query_snowflake20),
                 ("clique9", lambda: qgen.clique(9, 0x909)), ("cycle+tail", lambda: qgen.cycle_with_tail(5, 4, 17))]


@pytest.mark.parametrize("name,make", CONVERT_CASES, ids=[c[0] for c in CONVERT_CASES])
def test_conversion_gives_the_generators_query(name, make):
    g = make()
    info, keep = planner_info(spec_of(g))
    got = converted(info)
    assert (got["n"], got["nw"], got["n_ec"]) == (g.n, g.n_words, g.n_ec)
    for i in range(g.n):
        b = g.base[i]
        assert got["base"][i] == (b["rows"], b["tuples"], b["startup"], b["total"], b["width"]), i
    np.testing.assert_array_equal(got["edge_table"], g.edge_table)
    np.testing.assert_array_equal(got["moff"], g.ec_offsets)
    np.testing.assert_array_equal(got["members"], g.ec_members[:int(g.ec_offsets[-1])])
    assert got["sels"].tobytes() == g.ec_sels.tobytes()
    np.testing.assert_array_equal(got["fks"], g.ec_fks)
    assert got["vars"].tobytes() == g.ec_vars.tobytes()


def _wide_spec():
    """70 relations (two words per set): a 3-member class {0, 45, 66} with
    fks {45}, none, {0, 66}; a class {3, 69} whose relation 69 holds an FK
    to 3 and whose relation 3 has a unique index; the other relations have
    no edges (NULL edge sets)."""
    n = 70
    base = [(float(100 + i), float(200 + i), 0.5 * i, 10.0 + i, 4 + (i % 5)) for i in range(n)]
    edges = [set() for _ in range(n)]
    for a, b in ((0, 45), (0, 66), (45, 66), (3, 69)):
        edges[a].add(b)
        edges[b].add(a)
    z = dict(stadistinct=0.0, stanullfrac=0.0, mcvfreq=0.0, index_available=0, index_unique=0, index_rows=0.0,
             index_startup=0.0, index_total=0.0)
    va = [dict(z, stadistinct=-1.0, mcvfreq=0.25), dict(z, stadistinct=37.0, stanullfrac=0.125),
          dict(z, stadistinct=9.0, index_available=1, index_rows=3.0, index_startup=0.75, index_total=6.5)]
    vb = [dict(z, stadistinct=-1.0, index_available=1, index_unique=1, index_rows=1.0, index_startup=0.29,
               index_total=8.31), dict(z, stadistinct=1000.0)]
    classes = [([0, 45, 66], [0.5, 0.25, 0.125], [{45}, set(), {0, 66}], va),
               ([3, 69], [0.0625], [set(), {3}], vb)]
    return Spec(n, base, edges, classes), va + vb


def test_conversion_of_a_wide_query_with_a_three_member_class():
    spec, vs = _wide_spec()
    info, keep = planner_info(spec)
    got = converted(info)
    assert (got["n"], got["nw"], got["n_ec"]) == (70, 2, 2)
    assert got["moff"].tolist() == [0, 3, 5]
    assert got["members"].tolist() == [0, 45, 66, 3, 69]
    assert got["sels"].tolist() == [0.5, 0.25, 0.125, 0.0625]
    # relation r is bit r % 64 of word r // 64: {45}, {}, {0, 66}, {}, {3}
    assert got["fks"].tolist() == [[1 << 45, 0], [0, 0], [1, 1 << 2], [0, 0], [1 << 3, 0]]
    assert got["edge_table"][0].tolist() == [1 << 45, 1 << 2]
    assert got["edge_table"][45].tolist() == [1, 1 << 2]
    assert got["edge_table"][66].tolist() == [1 | (1 << 45), 0]
    assert got["edge_table"][69].tolist() == [1 << 3, 0]
    assert got["edge_table"][3].tolist() == [0, 1 << 5]
    assert not got["edge_table"][[1, 2, 4, 5, 6, 44, 46, 64, 65, 67, 68]].any()
    for j, v in enumerate(vs):
        for k, x in v.items():
            assert got["vars"][j][k] == np.float32(x), (j, k)
    assert got["base"][69] == (169.0, 269.0, 34.5, 79.0, 8)


def test_malformed_infos_are_refused():
    L = shim()
    spec, _ = _wide_spec()
    q = abi.Query()
    info, keep = planner_info(spec)
    info.n_eq_classes = 3                          # the list holds two
    assert L.gpuqo_amd_query_from_info(C.byref(info), C.byref(q)) == -1
    info.n_eq_classes = 1                          # ... and not one
    assert L.gpuqo_amd_query_from_info(C.byref(info), C.byref(q)) == -1
    info.n_eq_classes, info.n_eq_class_vars = 2, 4     # 5 members
    assert L.gpuqo_amd_query_from_info(C.byref(info), C.byref(q)) == -1
    info.n_eq_class_vars, info.n_eq_class_sels = 5, 3  # 4 sels
    assert L.gpuqo_amd_query_from_info(C.byref(info), C.byref(q)) == -1
    info.n_eq_class_sels, info.n_rels = 4, 0
    assert L.gpuqo_amd_query_from_info(C.byref(info), C.byref(q)) == -1
    info.n_rels = abi.GQ_MAX_RELS + 1
    assert L.gpuqo_amd_query_from_info(C.byref(info), C.byref(q)) == -1
    info.n_rels = 70
    assert L.gpuqo_amd_query_from_info(C.byref(info), C.byref(q)) == 0


# (global, ctype, gq_config field, a value other than the default)
GUCS = [("gpuqo_dpsub_bicc_enable", C.c_bool, "gpuqo_dpsub_bicc", True),
        ("gpuqo_dpsub_filter_enable", C.c_bool, "gpuqo_dpsub_filter", False),
        ("gpuqo_dpsub_filter_threshold", C.c_int, "gpuqo_dpsub_filter_threshold", 17),
        ("gpuqo_dpsub_filter_cpu_enum_threshold", C.c_int, "gpuqo_dpsub_filter_cpu_enum_threshold", 1025),
        ("gpuqo_dpsub_filter_keys_overprovisioning", C.c_int, "gpuqo_dpsub_filter_keys_overprovisioning", 101),
        ("gpuqo_dpsub_ccc_enable", C.c_bool, "gpuqo_dpsub_ccc", False),
        ("gpuqo_dpsub_csg_enable", C.c_bool, "gpuqo_dpsub_csg", False),
        ("gpuqo_dpsub_csg_threshold", C.c_int, "gpuqo_dpsub_csg_threshold", 129),
        ("gpuqo_dpsub_tree_enable", C.c_bool, "gpuqo_dpsub_tree", True),
        ("gpuqo_spanning_tree_enable", C.c_bool, "gpuqo_spanning_tree", True),
        ("gpuqo_n_parallel", C.c_int, "gpuqo_n_parallel", 10241),
        ("gpuqo_scratchpad_size_mb", C.c_int, "gpuqo_scratchpad_size_mb", 11),
        ("gpuqo_min_memo_size_mb", C.c_int, "gpuqo_min_memo_size_mb", 2),
        ("gpuqo_max_memo_size_mb", C.c_int, "gpuqo_max_memo_size_mb", 1025),
        ("gpuqo_idp_type", C.c_int, "gpuqo_idp_type", 2),
        ("gpuqo_idp_n_iters", C.c_int, "gpuqo_idp_n_iters", 9),
        ("gpuqo_dpe_n_threads", C.c_int, "gpuqo_dpe_n_threads", 5),
        ("gpuqo_cpu_dpsub_parallel_chunk_size", C.c_int, "gpuqo_cpu_dpsub_parallel_chunk_size", 65),
        ("gpuqo_dpe_pairs_per_depbuf", C.c_int, "gpuqo_dpe_pairs_per_depbuf", 129),
        ("seq_page_cost", C.c_double, "seq_page_cost", 1.5),
        ("random_page_cost", C.c_double, "random_page_cost", 4.5),
        ("cpu_tuple_cost", C.c_double, "cpu_tuple_cost", 0.02),
        ("cpu_index_tuple_cost", C.c_double, "cpu_index_tuple_cost", 0.006),
        ("cpu_operator_cost", C.c_double, "cpu_operator_cost", 0.003),
        ("disable_cost", C.c_double, "disable_cost", 2.0e10),
        ("effective_cache_size", C.c_int, "effective_cache_size", 524289),
        ("work_mem", C.c_int, "work_mem", 8192),
        ("enable_seqscan", C.c_bool, "enable_seqscan", False),
        ("enable_indexscan", C.c_bool, "enable_indexscan", False),
        ("enable_tidscan", C.c_bool, "enable_tidscan", False),
        ("enable_sort", C.c_bool, "enable_sort", False),
        ("enable_hashagg", C.c_bool, "enable_hashagg", False),
        ("enable_nestloop", C.c_bool, "enable_nestloop", False),
        ("enable_mergejoin", C.c_bool, "enable_mergejoin", False),
        ("enable_hashjoin", C.c_bool, "enable_hashjoin", False)]


def _cfg_dict(c):
    return {k: getattr(c, k) for k, _ in abi.Config._fields_}


def _from_gucs(algo):
    c = abi.Config()
    shim().gpuqo_amd_config_from_gucs(algo, C.byref(c))
    return _cfg_dict(c)


def test_default_gucs_give_the_default_config():
    want = abi.Config()
    assert gpuqo_amd.lib().gq_config_default(C.byref(want)) == 0
    want = _cfg_dict(want)
    for algo in (abi.ALGO["dpsize"], abi.ALGO["dpsub"]):
        want["gpuqo_algorithm"] = algo
        assert _from_gucs(algo) == want


@pytest.mark.parametrize("guc,ctype,field,value", GUCS, ids=[g[0] for g in GUCS])
def test_each_guc_sets_its_own_field(guc, ctype, field, value):
    var = ctype.in_dll(shim(), guc)
    old = var.value
    base = _from_gucs(1)
    try:
        var.value = value
        got = _from_gucs(1)
    finally:
        var.value = old
    changed = {k for k in got if got[k] != base[k]}
    assert changed == {field}
    assert got[field] == (int(value) if ctype is C.c_bool else value)


def test_every_config_field_has_a_guc():
    mapped = {g[2] for g in GUCS} | {"gpuqo_algorithm"}
    own = {"memo_kind", "device", "profile_kernels"}           # no GUC counterpart
    build_time = {"cost_function"}                             # a make variable in the reference
    assert mapped | own | build_time == {k for k, _ in abi.Config._fields_}


@pytest.mark.parametrize("define,want", [(None, 0), ("1", 1), ("2", 2)])
def test_cost_function_define(tmp_path, define, want):
    """The reference picks its cost function at build time (Makefile
    cost_function -> -DGPUQO_COST_FUNCTION=0|1|2); the shim built with the
    same define passes it as gq_config.cost_function."""
    out = str(tmp_path / "libshim_cf.so")
    lib_dir = os.path.join(HERE, "..", "postgres-gpuqo_amd", "lib")
    cmd = ["gcc", "-std=gnu99", "-O1", "-fPIC", "-shared", "-I" + os.path.join(SHIM_DIR, "include"),
           "-I" + os.path.join(HERE, "..", "include"), "-o", out,
           os.path.join(HERE, "..", "postgres-gpuqo_amd", "shim", "gpuqo_run_amd.c"),
           os.path.join(SHIM_DIR, "pg_standin.c"), "-L" + lib_dir, "-lgpuqo_amd", "-Wl,-rpath," + lib_dir]
    if define is not None:
        cmd.insert(1, "-DGPUQO_COST_FUNCTION=" + define)
    subprocess.run(cmd, check=True)
    gpuqo_amd.lib()
    L = C.CDLL(out)
    c = abi.Config()
    L.gpuqo_amd_config_from_gucs(abi.ALGO["dpsub"], C.byref(c))
    assert c.cost_function == want


def members_of(b):
    L, out, x = shim(), [], -1
    while True:
        x = L.bms_next_member(b, x)
        if x < 0:
            return out
        out.append(x)


def tree_nodes(t):
    """Pre-order walk of a QueryTreeC: (members, rows, startup, total, width,
    has_left, has_right)."""
    out = []

    def walk(p):
        n = p.contents
        out.append((members_of(n.id), n.rows, n.cost.startup, n.cost.total, n.width, bool(n.left), bool(n.right)))
        if n.left:
            walk(n.left)
        if n.right:
            walk(n.right)
    walk(t)
    return out


def _plan_out(nodes):
    arr = (abi.PlanNode * len(nodes))()
    for i, (r, rows, st, tot, w, l, rr) in enumerate(nodes):
        arr[i].relids = (C.c_uint64 * abi.GQ_MAX_WORDS)(*r)
        arr[i].rows, arr[i].cost.startup, arr[i].cost.total = rows, st, tot
        arr[i].width, arr[i].left, arr[i].right = w, l, rr
    return abi.PlanOut(C.cast(arr, C.POINTER(abi.PlanNode)), len(nodes), len(nodes)), arr


def expected_tree(nodes):
    """The pre-order tree a plan stands for, relation r as member r + 1."""
    out = []

    def walk(i):
        r, rows, st, tot, w, l, rr = nodes[i]
        mem = [b + 1 for b in range(64 * abi.GQ_MAX_WORDS) if (r[b // 64] >> (b % 64)) & 1]
        out.append((mem, rows, st, tot, w, l >= 0, rr >= 0))
        if l >= 0:
            walk(l)
        if rr >= 0:
            walk(rr)
    walk(0)
    return out


@pytest.mark.parametrize("name,make", CONVERT_CASES[:2], ids=[c[0] for c in CONVERT_CASES[:2]])
def test_tree_from_an_oracle_plan(name, make):
    g = make()
    st, nodes, _ = oracle.plan(g, abi.default_config(algorithm="cpu_dpsub_bicc"))
    assert st == 0
    out, keep = _plan_out(nodes)
    before = shim().pg_standin_palloc_count()
    t = shim().gpuqo_amd_tree_from_plan(C.byref(out))
    got = tree_nodes(t)
    assert got == expected_tree(nodes)
    assert got[0][0] == list(range(1, g.n + 1))
    leaves = [x for x in got if not x[5]]
    assert len(got) == 2 * g.n - 1 and len(leaves) == g.n
    assert all(not x[6] and len(x[0]) == 1 for x in leaves)
    # every node and every id in the caller's memory context
    assert shim().pg_standin_palloc_count() - before >= len(got)


def test_tree_from_an_empty_plan_is_null():
    out = abi.PlanOut(None, 0, 0)
    assert not shim().gpuqo_amd_tree_from_plan(C.byref(out))


def test_gpuqo_run_cpu_algorithm_returns_null_with_a_warning():
    """gq_plan refuses a CPU algorithm before any device work: the shim
    returns NULL (the reference's failure value) after a WARNING that
    carries the library's message."""
    L = shim()
    info, keep = planner_info(spec_of(qgen.query_star12()))
    assert not L.gpuqo_run(abi.ALGO["cpu_dpsub"], C.byref(info))
    lvl = C.c_int()
    msg = L.pg_standin_last_message(C.byref(lvl))
    assert lvl.value == 19 and msg.startswith(b"gpuqo_amd: ") and b"not a GPU algorithm" in msg
    assert L.pg_standin_interrupt_checks() == 0


def test_gpuqo_run_malformed_info_returns_null():
    L = shim()
    spec, _ = _wide_spec()
    info, keep = planner_info(spec)
    info.n_eq_classes = 5
    assert not L.gpuqo_run(abi.ALGO["dpsub"], C.byref(info))
    assert L.pg_standin_last_message(None) == b"gpuqo_amd: malformed planner info"


# ------------------------------------------------------------------ GPU

def _guc(name, ctype):
    return ctype.in_dll(shim(), name)


RUN_CASES = [("star12", qgen.query_star12), ("snowflake20", qgen.query_snowflake20)]
ALGOS = [("dpsub", {}), ("mpdp", {"gpuqo_dpsub_bicc_enable": True}), ("dpsize", {}),
         ("uniondp", {"gpuqo_idp_type": 2, "gpuqo_idp_n_iters": 8})]


@pytest.mark.gpu
@pytest.mark.parametrize("algo,gucs", ALGOS, ids=[a[0] for a in ALGOS])
@pytest.mark.parametrize("name,make", RUN_CASES, ids=[c[0] for c in RUN_CASES])
def test_gpu_gpuqo_run_equals_gq_plan(name, make, algo, gucs):
    L = shim()
    g = make()
    code = abi.ALGO["dpsize" if algo == "dpsize" else "dpsub"]
    types = {"gpuqo_dpsub_bicc_enable": C.c_bool, "gpuqo_idp_type": C.c_int, "gpuqo_idp_n_iters": C.c_int}
    saved = {k: _guc(k, types[k]).value for k in gucs}
    try:
        for k, v in gucs.items():
            _guc(k, types[k]).value = v
        cfg = abi.Config()
        L.gpuqo_amd_config_from_gucs(code, C.byref(cfg))
        want, _ = gpuqo_amd.plan(g, cfg)
        info, keep = planner_info(spec_of(g))
        t = L.gpuqo_run(code, C.byref(info))
    finally:
        for k, v in saved.items():
            _guc(k, types[k]).value = v
    assert t, L.pg_standin_last_message(None)
    got = tree_nodes(t)
    exp = expected_tree(want)
    assert len(got) == len(exp) == 2 * g.n - 1
    fb = gpuqo_amd.fbits
    assert_same_plan([(m, fb(a), fb(b), fb(c), w, l, r) for (m, a, b, c, w, l, r) in got],
                     [(m, fb(a), fb(b), fb(c), w, l, r) for (m, a, b, c, w, l, r) in exp], f"{name} {algo}")


@pytest.mark.gpu
def test_gpu_gpuqo_run_interrupted_returns_null_after_check_for_interrupts():
    L = shim()
    info, keep = planner_info(spec_of(qgen.query_star12()))
    gpuqo_amd.lib().gq_interrupt(1)
    try:
        t = L.gpuqo_run(abi.ALGO["dpsub"], C.byref(info))
    finally:
        gpuqo_amd.lib().gq_interrupt(0)
    assert not t
    assert L.pg_standin_interrupt_checks() == 1
    assert b"interrupted" in L.pg_standin_last_message(None)
    assert L.gpuqo_run(abi.ALGO["dpsub"], C.byref(info))
