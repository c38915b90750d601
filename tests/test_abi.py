"""The C ABI library loads, exports every entry point include/*.h declares,
and its structures match the Python mirror.  No device work is done here."""
import ctypes as C
import glob
import os
import re

import pytest

import gpuqo_amd
from gpuqo_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(gq_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for f in ("gq_plan", "gq_plan_batch", "gq_config_default", "gq_server_create", "gq_server_submit",
              "gq_server_submit_cfg", "gq_server_poll", "gq_server_destroy", "gq_last_error", "gq_init"):
        assert f in names


def test_library_exports_all_declared_symbols():
    lib = gpuqo_amd.lib()
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_struct_sizes():
    assert C.sizeof(abi.BaseRel) == 20
    assert C.sizeof(abi.VarInfo) == 28
    assert C.sizeof(abi.PlanNode) == 56
    assert C.sizeof(abi.Query) == 72


def test_config_default_matches_mirror():
    lib = gpuqo_amd.lib()
    c = abi.Config()
    assert lib.gq_config_default(C.byref(c)) == 0
    d = abi.default_config()
    for name, _ in abi.Config._fields_:
        assert getattr(c, name) == getattr(d, name), name


def test_version_string():
    assert b"gfx950" in gpuqo_amd.lib().gq_version()


def test_widths_are_validated_before_any_device_call():
    """Join widths are int sums of base widths: a negative base width, or
    base widths summing past 2^31-1, is rejected as invalid input."""
    from gpuqo_amd import qgen
    cfg = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
    g = qgen.chain(4, 1)
    g.build()
    g.base["width"][2] = -1
    with pytest.raises(gpuqo_amd.GpuqoError) as e:
        gpuqo_amd.plan(g, cfg)
    assert e.value.status == abi.GQ_ERR_INVALID and "negative width" in str(e.value)
    g = qgen.chain(4, 1)
    g.build()
    g.base["width"][:] = 0x40000000
    with pytest.raises(gpuqo_amd.GpuqoError) as e:
        gpuqo_amd.plan(g, cfg)
    assert e.value.status == abi.GQ_ERR_INVALID and "widths sum" in str(e.value)


def test_tuning_defaults_and_validation():
    """The launch-shape knobs start at their defaults; values outside the
    accepted ranges are refused and leave the knobs unchanged."""
    d = gpuqo_amd.tuning_get()
    assert d == {"tree_two_lane_waves": 4096, "general_grid_max": 512, "tree_task_iters": 16,
                 "tree_expand_max_rels": 20, "tree_one_xcd_wgs": 0}
    for bad in ({"tree_task_iters": 0}, {"tree_task_iters": 257}, {"general_grid_max": 0},
                {"tree_two_lane_waves": -1}, {"tree_expand_max_rels": 33}, {"tree_one_xcd_wgs": 257},
                {"tree_one_xcd_wgs": -1}):
        with pytest.raises(gpuqo_amd.GpuqoError) as e:
            gpuqo_amd.tuning_set(**bad)
        assert e.value.status == abi.GQ_ERR_INVALID
        assert gpuqo_amd.tuning_get() == d
    gpuqo_amd.tuning_set(tree_task_iters=8)
    assert gpuqo_amd.tuning_get()["tree_task_iters"] == 8
    gpuqo_amd.tuning_set(**d)
    assert gpuqo_amd.tuning_get() == d
