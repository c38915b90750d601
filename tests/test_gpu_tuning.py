"""Launch-shape knobs change no plan: every query plans to the same join
tree and bits under extreme knob values and after the autotuner ran."""
import pytest

import gpuqo_amd
from helpers import assert_same_plan, cfg, gpu_bits
from gpuqo_amd import qgen

pytestmark = pytest.mark.gpu

DEFAULTS = {"tree_two_lane_waves": 4096, "general_grid_max": 512, "tree_task_iters": 16,
            "tree_expand_max_rels": 20, "tree_one_xcd_wgs": 0}


def _sample():
    return [qgen.snowflake(20, 0x720), qgen.star(17, 0x717), qgen.chain(14, 0x714), qgen.cycle(16, 0x716),
            qgen.snowflake(22, 0x722), qgen.clique(9, 0x709)]


def _plans(qs, c):
    single = [gpu_bits(g, c)[0] for g in qs]
    batch = [gpuqo_amd.plan_bits(n) for n, _ in gpuqo_amd.plan_batch(qs, c)]
    return single, batch


@pytest.mark.parametrize("knobs", [{"tree_two_lane_waves": 1}, {"tree_two_lane_waves": 1 << 24},
                                   {"tree_task_iters": 1, "general_grid_max": 1},
                                   {"tree_task_iters": 64, "general_grid_max": 4096},
                                   {"tree_expand_max_rels": 0}, {"tree_expand_max_rels": 32},
                                   {"tree_one_xcd_wgs": 32}, {"tree_one_xcd_wgs": 256}])
def test_knobs_change_no_plan(knobs):
    qs = _sample()
    for bicc in (1, 0):
        c = cfg("dpsub", bicc)
        gpuqo_amd.tuning_set(**DEFAULTS)
        want_s, want_b = _plans(qs, c)
        try:
            gpuqo_amd.tuning_set(**knobs)
            got_s, got_b = _plans(qs, c)
        finally:
            gpuqo_amd.tuning_set(**DEFAULTS)
        for g, a, b in zip(qs, got_s, want_s):
            assert_same_plan(a, b, f"{g.shape}{g.n} single {knobs}")
        for g, a, b in zip(qs, got_b, want_b):
            assert_same_plan(a, b, f"{g.shape}{g.n} batch {knobs}")


def test_autotune_keeps_plans():
    qs = _sample()
    c = cfg("dpsub", 1)
    gpuqo_amd.tuning_set(**DEFAULTS)
    want_s, want_b = _plans(qs, c)
    try:
        t = gpuqo_amd.autotune(qs, c, reps=1)
        assert t == gpuqo_amd.tuning_get()
        assert t["tree_two_lane_waves"] in (1024, 2048, 4096, 8192, 16384)
        assert t["tree_expand_max_rels"] in (0, 16, 18, 20, 22)
        assert t["tree_one_xcd_wgs"] in (0, 8, 32, 64)
        assert t["tree_task_iters"] in (4, 8, 16, 32)
        assert t["general_grid_max"] in (256, 512, 1024, 2048)
        got_s, got_b = _plans(qs, c)
    finally:
        gpuqo_amd.tuning_set(**DEFAULTS)
    for g, a, b in zip(qs, got_s, want_s):
        assert_same_plan(a, b, f"{g.shape}{g.n} single after autotune")
    for g, a, b in zip(qs, got_b, want_b):
        assert_same_plan(a, b, f"{g.shape}{g.n} batch after autotune")


def test_autotune_groups_leave_other_knobs():
    qs = _sample()[:3]
    c = cfg("dpsub", 1)
    try:
        gpuqo_amd.tuning_set(**DEFAULTS)
        t = gpuqo_amd.autotune(qs, c, reps=1, knobs=gpuqo_amd.TUNE_BATCH)
        assert t["tree_two_lane_waves"] == DEFAULTS["tree_two_lane_waves"]
        assert t["tree_expand_max_rels"] == DEFAULTS["tree_expand_max_rels"]
        assert t["tree_one_xcd_wgs"] == DEFAULTS["tree_one_xcd_wgs"]
        gpuqo_amd.tuning_set(**DEFAULTS)
        t = gpuqo_amd.autotune(qs, c, reps=1, knobs=gpuqo_amd.TUNE_SINGLE)
        assert t["tree_task_iters"] == DEFAULTS["tree_task_iters"]
        assert t["general_grid_max"] == DEFAULTS["general_grid_max"]
    finally:
        gpuqo_amd.tuning_set(**DEFAULTS)
