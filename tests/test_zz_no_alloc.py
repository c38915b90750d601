"""gq_plan_noalloc: the allocation-free variant of gq_plan (a static device
array holds the DP).  Through the C ABI, its plans equal gq_plan's bit for
bit, with the same pair and set counts, on two calls in a row with the
second query smaller -- for tree DPs (the fused start/walk kernels), for
general graphs (2^n filter, LDS pair queue) and for MPDP sets with large
biconnected blocks (the big-set list) -- and the calls allocate no device
memory.  Limits are rejected before any device work."""
import ctypes as C

import pytest

import gpuqo_amd
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan


def _noalloc(g, c):
    L = gpuqo_amd.lib()
    L.gq_plan_noalloc.argtypes = [C.POINTER(abi.Query), C.POINTER(abi.Config), C.POINTER(abi.PlanOut)]
    cap = 2 * g.n
    nodes = (abi.PlanNode * cap)()
    out = abi.PlanOut(C.cast(nodes, C.POINTER(abi.PlanNode)), cap, 0)
    st = L.gq_plan_noalloc(C.byref(g.to_c()), C.byref(c), C.byref(out))
    return st, gpuqo_amd.plan_bits(gpuqo_amd.plan_tuples(out)), out.stats


def _free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


MPDP = dict(algorithm="dpsub", dpsub_bicc=1)
DPSUB = dict(algorithm="dpsub", dpsub_bicc=0)
PAIRS = [
    ("snowflake20-then-star12", (qgen.query_snowflake20(), MPDP), (qgen.query_star12(), DPSUB)),
    ("clique14-then-cycle10", (qgen.clique(14, 0x7701), DPSUB), (qgen.cycle(10, 0x7702), MPDP)),
    ("chorded17-then-chain9", (qgen.chorded_cycle(17, 5, 0x7703), MPDP), (qgen.chain(9, 0x7704), DPSUB)),
]


@pytest.mark.gpu
def test_noalloc_allocates_no_device_memory():
    L = gpuqo_amd.lib()
    assert L.gq_init(0) == 0
    g = qgen.query_snowflake20()
    c = abi.default_config(**MPDP)
    before = _free_device_bytes()
    for _ in range(3):
        st, _, _ = _noalloc(g, c)
        assert st == 0
    after = _free_device_bytes()
    assert after >= before - (1 << 20), (before, after)


@pytest.mark.gpu
@pytest.mark.parametrize("name,first,second", PAIRS, ids=[p[0] for p in PAIRS])
def test_noalloc_matches_gq_plan_on_two_calls(name, first, second):
    L = gpuqo_amd.lib()
    assert L.gq_init(0) == 0
    for g, kw in (first, second):                     # the second call has the smaller query
        c = abi.default_config(**kw)
        st, got, stats = _noalloc(g, c)
        assert st == 0, L.gq_last_error()
        want, wst = gpuqo_amd.plan(g, c)
        assert_same_plan(got, gpuqo_amd.plan_bits(want), f"{name} n={g.n}")
        assert stats.pairs == wst.pairs and stats.sets == wst.sets
    # and once more after gq_plan has used its own buffers
    g, kw = first
    st, got, _ = _noalloc(g, abi.default_config(**kw))
    assert st == 0
    assert_same_plan(got, gpuqo_amd.plan_bits(gpuqo_amd.plan(g, abi.default_config(**kw))[0]), name)


def test_noalloc_rejects_what_it_cannot_plan():
    """Checked before any device work (runs without a GPU)."""
    L = gpuqo_amd.lib()
    g21 = qgen.snowflake(21, 0x7705)
    assert _noalloc(g21, abi.default_config(**MPDP))[0] == abi.GQ_ERR_UNSUPPORTED
    assert b"at most 20" in L.gq_last_error()
    g = qgen.star(12, 0x7706)
    assert _noalloc(g, abi.default_config(algorithm="dpsize"))[0] == abi.GQ_ERR_UNSUPPORTED
    idp = abi.default_config(algorithm="dpsub", dpsub_bicc=1, idp_type="idp2", idp_n_iters=5)
    assert _noalloc(g, idp)[0] == abi.GQ_ERR_UNSUPPORTED
    assert _noalloc(g, abi.default_config(algorithm="dpsub", memo_kind=3))[0] == abi.GQ_ERR_UNSUPPORTED


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["simple", "cout"])
def test_noalloc_cost_functions(fn):
    """The allocation-free path plans under the simple and C_out cost
    functions as gq_plan does."""
    for g, kw in ((qgen.query_snowflake20(), MPDP), (qgen.clique(12, 0x7705), DPSUB)):
        c = abi.default_config(cost_function=fn, **kw)
        st, got, _ = _noalloc(g, c)
        assert st == 0
        want, _ = gpuqo_amd.plan(g, c)
        assert_same_plan(got, gpuqo_amd.plan_bits(want), f"noalloc {fn}")
