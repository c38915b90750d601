"""Failure handling at the C ABI, on the CPU: no C++ exception crosses the
boundary.  A host allocation that fails inside gq_plan must come back as
GQ_ERR_NOMEM with a message, and the process must live on -- a
std::bad_alloc reaching std::terminate would end the Postgres backend that
calls the planner.

The failure is real, not simulated: a child process lowers its address-
space limit (RLIMIT_AS) just above what it already maps and then asks
gq_plan to convert a 250-relation clique (31,125 classes: megabytes of host
vectors).  The query is converted on the host before any device call, so
this needs no GPU."""
import os
import subprocess
import sys
import textwrap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = textwrap.dedent(r"""
    import ctypes as C, resource, sys
    sys.path[:0] = [{pkg!r}]
    import gpuqo_amd
    from gpuqo_amd import abi, qgen
    L = gpuqo_amd.lib()
    g = qgen.clique(250, 7)
    q = g.to_c()
    cfg = abi.default_config(algorithm="dpsub", dpsub_bicc=1, idp_type="union", idp_n_iters=25)
    nodes = (abi.PlanNode * 500)()
    out = abi.PlanOut(C.cast(nodes, C.POINTER(abi.PlanNode)), 500, 0)
    vm = 0
    for line in open("/proc/self/status"):
        if line.startswith("VmSize:"):
            vm = int(line.split()[1]) * 1024
    soft, hard = resource.getrlimit(resource.RLIMIT_AS)
    resource.setrlimit(resource.RLIMIT_AS, (vm + (1 << 20), hard))
    st = L.gq_plan(C.byref(q), C.byref(cfg), C.byref(out))
    resource.setrlimit(resource.RLIMIT_AS, (soft, hard))
    msg = L.gq_last_error().decode()
    print("STATUS", st, msg, flush=True)
""")


def test_host_allocation_failure_returns_nomem_and_process_survives():
    code = CHILD.format(pkg=os.path.join(ROOT, "postgres-gpuqo_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    line = [x for x in r.stdout.splitlines() if x.startswith("STATUS")][-1]
    status = int(line.split()[1])
    assert status == 3, line                    # GQ_ERR_NOMEM
    assert "allocation failed" in line


def test_status_codes_match_the_header():
    import re
    from gpuqo_amd import abi
    txt = open(os.path.join(ROOT, "include", "gpuqo_amd.h")).read()
    codes = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"\b(GQ_(?:OK|ERR_\w+))\s*=\s*(\d+)", txt))
    for name, v in codes.items():
        assert getattr(abi, name) == v, name
    assert codes["GQ_ERR_INTERNAL"] == 9
