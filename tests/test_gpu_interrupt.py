"""gq_interrupt inside a running DP (the reference's CHECK_FOR_INTERRUPTS,
polled per level: gpuqo_dpsub_common.cu:300, gpuqo_dpsize.cu:451).

A host thread raises the flag while gq_plan runs an 18-relation clique DP
(tens of milliseconds on the device).  The planner stops enqueuing levels,
lets the enqueued ones finish (every kernel exits normally) and returns
GQ_ERR_INTERRUPTED; with the flag cleared, the next call plans the same
query bit-exactly.  The flag is forced from the host only."""
import ctypes as C
import threading
import time

import pytest

import gpuqo_amd
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan


def _plan_status(g, c, box):
    cap = 2 * g.n
    nodes = (abi.PlanNode * cap)()
    out = abi.PlanOut(C.cast(nodes, C.POINTER(abi.PlanNode)), cap, 0)
    box["st"] = gpuqo_amd.lib().gq_plan(C.byref(g.to_c()), C.byref(c), C.byref(out))
    box["msg"] = gpuqo_amd.lib().gq_last_error()       # the message is per thread
    box["nodes"] = gpuqo_amd.plan_tuples(out)


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["dpsub", "dpsize"])
def test_interrupt_during_clique18_then_clean_plan(algo):
    L = gpuqo_amd.lib()
    g = qgen.query_clique18()
    c = abi.default_config(algorithm=algo, dpsub_bicc=0)
    L.gq_interrupt(0)
    want, st0 = gpuqo_amd.plan(g, c)              # warm: buffers allocated, reference plan
    box = {}
    try:
        t = threading.Thread(target=_plan_status, args=(g, c, box))
        t0 = time.perf_counter()
        t.start()
        time.sleep(0.25 * st0.gpu_ms / 1000.0)    # about a quarter of the DP
        L.gq_interrupt(1)
        t.join(timeout=60)
        dt = time.perf_counter() - t0
        assert not t.is_alive()
    finally:
        L.gq_interrupt(0)
    assert box["st"] == abi.GQ_ERR_INTERRUPTED, (box["st"], box["msg"])
    assert b"interrupted" in box["msg"]
    print(f"{algo}: interrupted after {dt * 1e3:.1f} ms of a {st0.gpu_ms:.1f} ms DP")
    again, _ = gpuqo_amd.plan(g, c)
    assert_same_plan(gpuqo_amd.plan_bits(again), gpuqo_amd.plan_bits(want), f"clique18-{algo} after interrupt")


@pytest.mark.gpu
def test_interrupt_flag_set_before_the_call():
    L = gpuqo_amd.lib()
    g = qgen.clique(16, 5)
    c = abi.default_config(algorithm="dpsub", dpsub_bicc=0)
    L.gq_interrupt(1)
    try:
        box = {}
        _plan_status(g, c, box)
    finally:
        L.gq_interrupt(0)
    assert box["st"] == abi.GQ_ERR_INTERRUPTED
    nodes, _ = gpuqo_amd.plan(g, c)
    assert len(nodes) == 2 * g.n - 1
