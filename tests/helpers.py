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
