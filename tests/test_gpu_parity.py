"""GPU DP vs the CPU oracle: same join tree, same rows/costs bit for bit."""
import pytest

from helpers import assert_same_plan, cfg, gpu_bits, oracle_bits, small_graphs
from gpuqo_amd import qgen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,g", small_graphs(), ids=lambda x: x if isinstance(x, str) else "")
@pytest.mark.parametrize("bicc", [0, 1])
def test_small_graphs(name, g, bicc):
    c = cfg("dpsub", bicc)
    want, _ = oracle_bits(g, c)
    got, _ = gpu_bits(g, c)
    assert_same_plan(got, want, f"{name} bicc={bicc}")


def test_config1_star12_dpsub():
    g = qgen.config1()
    c = cfg("dpsub", 0)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, "star-12 dpsub")
    assert gst.pairs == ost.pairs == 11 * 2 ** 11


def test_config2_snowflake20_mpdp():
    g = qgen.config2()
    c = cfg("dpsub", 1)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, "snowflake-20 mpdp")
    assert gst.pairs == ost.pairs
    assert gst.sets == ost.sets


def test_clique14_dpsub():
    g = qgen.clique(14, 0x77)
    c = cfg("dpsub", 0)
    want, ost = oracle_bits(g, c)
    got, gst = gpu_bits(g, c)
    assert_same_plan(got, want, "clique-14 dpsub")
    assert gst.pairs == 3 ** 14 - 2 ** 15 + 1


@pytest.mark.parametrize("seed", range(4))
def test_mixed_batch_members(seed):
    for g in qgen.mixed_batch(6, 1000 + 97 * seed, nmin=8, nmax=16):
        c = cfg("dpsub", 1)
        want, _ = oracle_bits(g, c)
        got, _ = gpu_bits(g, c)
        assert_same_plan(got, want, f"{g.shape}{g.n}")
