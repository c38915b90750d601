"""Host-side Python binding of the MI355X join-order optimizer core.

The compute path is the C ABI library lib/libgpuqo_amd.so (HIP kernels for
gfx950 plus the C++ host drivers).  This module only marshals arguments; it
has no fallback: if the library is missing, calls raise.
"""
import ctypes as C
import os
import struct

from . import abi, qgen  # noqa: F401

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libgpuqo_amd.so")
_lib = None


class GpuqoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"gpuqo status {status}: {msg}")
        self.status = status


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.gq_config_default.argtypes = [C.POINTER(abi.Config)]
        L.gq_init.argtypes = [C.c_int32]
        L.gq_plan.argtypes = [C.POINTER(abi.Query), C.POINTER(abi.Config), C.POINTER(abi.PlanOut)]
        L.gq_plan_batch.argtypes = [C.c_int32, C.POINTER(C.POINTER(abi.Query)), C.POINTER(abi.Config),
                                    C.POINTER(abi.PlanOut)]
        L.gq_last_error.restype = C.c_char_p
        L.gq_version.restype = C.c_char_p
        L.gq_interrupt.argtypes = [C.c_int32]
        L.gq_interrupt.restype = None
        L.gq_server_create.argtypes = [C.POINTER(abi.Config), C.c_uint64, C.c_int32, C.POINTER(C.c_void_p)]
        L.gq_server_submit.argtypes = [C.c_void_p, C.POINTER(abi.Query), C.c_uint64]
        L.gq_server_submit_cfg.argtypes = [C.c_void_p, C.POINTER(abi.Query), C.POINTER(abi.Config), C.c_uint64]
        L.gq_server_poll.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_int32), C.POINTER(abi.PlanOut)]
        L.gq_server_destroy.argtypes = [C.c_void_p]
        L.gq_tuning_get.argtypes = [C.POINTER(abi.Tuning)]
        L.gq_tuning_set.argtypes = [C.POINTER(abi.Tuning)]
        L.gq_autotune.argtypes = [C.c_int32, C.POINTER(C.POINTER(abi.Query)), C.POINTER(abi.Config), C.c_int32,
                                  C.c_int32, C.POINTER(abi.Tuning)]
        L.gq_plan_snapshot.argtypes = [C.POINTER(abi.Query), C.POINTER(abi.Config), C.POINTER(abi.PlanOut),
                                       C.POINTER(C.c_void_p)]
        L.gq_snapshot_subplan.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(abi.PlanOut)]
        L.gq_snapshot_info.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]
        L.gq_snapshot_export.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]
        L.gq_snapshot_import.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]
        L.gq_snapshot_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        msg = lib().gq_last_error()
        raise GpuqoError(st, msg.decode() if msg else "")


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def plan_tuples(out):
    """Plan nodes as tuples (relids tuple, rows, startup, total, width, left, right)."""
    res = []
    for i in range(out.n_nodes):
        n = out.nodes[i]
        res.append((tuple(n.relids), n.rows, n.cost.startup, n.cost.total, n.width, n.left, n.right))
    return res


def plan_bits(nodes):
    """Same tuples with floats replaced by their IEEE-754 bit patterns."""
    return [(r, fbits(a), fbits(b), fbits(c), w, l, rr) for (r, a, b, c, w, l, rr) in nodes]


def default_config(**kw):
    return abi.default_config(**kw)


def plan(query, cfg):
    """Plan one query on the GPU; returns (nodes, stats)."""
    n = query.n
    cap = 2 * n
    nodes = (abi.PlanNode * cap)()
    out = abi.PlanOut(C.cast(nodes, C.POINTER(abi.PlanNode)), cap, 0)
    _check(lib().gq_plan(C.byref(query.to_c()), C.byref(cfg), C.byref(out)))
    return plan_tuples(out), out.stats


def plan_raw(query, cfg):
    """gq_plan without converting the result (for timing the C ABI call);
    returns the gq_plan_out and its node buffer."""
    cap = 2 * query.n
    nodes = (abi.PlanNode * cap)()
    out = abi.PlanOut(C.cast(nodes, C.POINTER(abi.PlanNode)), cap, 0)
    _check(lib().gq_plan(C.byref(query.to_c()), C.byref(cfg), C.byref(out)))
    return out, nodes


class BatchArgs:
    """Prebuilt gq_plan_batch arguments for a fixed list of queries."""

    def __init__(self, queries):
        k = len(queries)
        self.k = k
        self.qptrs = (C.POINTER(abi.Query) * k)(*[C.pointer(q.to_c()) for q in queries])
        self.bufs = [(abi.PlanNode * (2 * q.n))() for q in queries]
        self.outs = (abi.PlanOut * k)()
        for i, q in enumerate(queries):
            self.outs[i].nodes = C.cast(self.bufs[i], C.POINTER(abi.PlanNode))
            self.outs[i].capacity = 2 * q.n

    def run(self, cfg):
        _check(lib().gq_plan_batch(self.k, self.qptrs, C.byref(cfg), self.outs))
        return self.outs


def plan_batch(queries, cfg):
    """Plan many queries in one batched device pass; returns [(nodes, stats)]."""
    k = len(queries)
    qptrs = (C.POINTER(abi.Query) * k)(*[C.pointer(q.to_c()) for q in queries])
    bufs = [(abi.PlanNode * (2 * q.n))() for q in queries]
    outs = (abi.PlanOut * k)()
    for i, q in enumerate(queries):
        outs[i].nodes = C.cast(bufs[i], C.POINTER(abi.PlanNode))
        outs[i].capacity = 2 * q.n
    _check(lib().gq_plan_batch(k, qptrs, C.byref(cfg), outs))
    return [(plan_tuples(outs[i]), outs[i].stats) for i in range(k)]


class Snapshot:
    """A device-resident memo snapshot (gq_plan_snapshot / gq_snapshot_*):
    the plan of the query, then the plans of its connected sub-queries."""

    def __init__(self, query=None, cfg=None, handle=None):
        self.h = None
        self.plan, self.stats = None, None
        if handle is not None:
            self.h = handle
            return
        cap = 2 * query.n
        nodes = (abi.PlanNode * cap)()
        out = abi.PlanOut(C.cast(nodes, C.POINTER(abi.PlanNode)), cap, 0)
        h = C.c_void_p()
        _check(lib().gq_plan_snapshot(C.byref(query.to_c()), C.byref(cfg), C.byref(out), C.byref(h)))
        self.h = h
        self.plan, self.stats = plan_tuples(out), out.stats

    def subplan(self, relids):
        """Plan nodes of the connected sub-query `relids` (a relation bit mask)."""
        n_rels, _sets, _bytes = self.info()
        cap = 2 * n_rels
        nodes = (abi.PlanNode * cap)()
        out = abi.PlanOut(C.cast(nodes, C.POINTER(abi.PlanNode)), cap, 0)
        words = (C.c_uint64 * abi.GQ_MAX_WORDS)(relids)
        _check(lib().gq_snapshot_subplan(self.h, words, C.byref(out)))
        return plan_tuples(out)

    def info(self):
        n, s, b = C.c_int32(), C.c_uint64(), C.c_uint64()
        _check(lib().gq_snapshot_info(self.h, C.byref(n), C.byref(s), C.byref(b)))
        return n.value, s.value, b.value

    def export(self):
        size = C.c_uint64()
        _check(lib().gq_snapshot_export(self.h, None, 0, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        _check(lib().gq_snapshot_export(self.h, buf, size.value, C.byref(size)))
        return buf.raw[:size.value]

    @staticmethod
    def load(data):
        h = C.c_void_p()
        _check(lib().gq_snapshot_import(data, len(data), C.byref(h)))
        return Snapshot(handle=h)

    def close(self):
        if self.h:
            _check(lib().gq_snapshot_free(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """The continuous-batching planning server (gq_server_*)."""

    def __init__(self, cfg, memo_budget_bytes, max_inflight=8):
        self.cfg = cfg
        h = C.c_void_p()
        _check(lib().gq_server_create(C.byref(cfg), memo_budget_bytes, max_inflight, C.byref(h)))
        self.h = h
        cap = 2 * abi.GQ_MAX_RELS
        self._buf = (abi.PlanNode * cap)()
        self._out = abi.PlanOut(C.cast(self._buf, C.POINTER(abi.PlanNode)), cap, 0)

    def submit(self, query, tag, cfg=None):
        """Queue one query (copied); `cfg` overrides the server's config."""
        if cfg is None:
            _check(lib().gq_server_submit(self.h, C.byref(query.to_c()), tag))
        else:
            _check(lib().gq_server_submit_cfg(self.h, C.byref(query.to_c()), C.byref(cfg), tag))

    def poll(self, wait_ms=-1):
        """One finished request as (tag, status, nodes, stats), or None."""
        done, tag, st = C.c_int32(0), C.c_uint64(0), C.c_int32(0)
        _check(lib().gq_server_poll(self.h, wait_ms, C.byref(done), C.byref(tag), C.byref(st),
                                    C.byref(self._out)))
        if not done.value:
            return None
        stats = abi.PlanStats()
        C.pointer(stats)[0] = self._out.stats
        return tag.value, st.value, plan_tuples(self._out), stats

    def close(self):
        if self.h:
            _check(lib().gq_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _tuning_dict(t):
    return {name: getattr(t, name) for name, _ in abi.Tuning._fields_}


def tuning_get():
    """The launch-shape knobs in force (gq_tuning) as a dict."""
    t = abi.Tuning()
    _check(lib().gq_tuning_get(C.byref(t)))
    return _tuning_dict(t)


def tuning_set(**kw):
    """Set some launch-shape knobs (the others keep their values)."""
    t = abi.Tuning()
    _check(lib().gq_tuning_get(C.byref(t)))
    names = {f for f, _ in abi.Tuning._fields_}
    for k, v in kw.items():
        if k not in names:
            raise KeyError(k)
        setattr(t, k, v)
    _check(lib().gq_tuning_set(C.byref(t)))


TUNE_SINGLE, TUNE_BATCH = 1, 2


def autotune(queries, cfg, reps=3, knobs=TUNE_SINGLE | TUNE_BATCH):
    """Time the sample queries under candidate knob values (gq_autotune) and
    keep the fastest; returns all knobs after tuning."""
    cq = [q.to_c() for q in queries]
    arr = (C.POINTER(abi.Query) * len(cq))(*[C.pointer(q) for q in cq])
    t = abi.Tuning()
    _check(lib().gq_autotune(len(cq), arr, C.byref(cfg), reps, knobs, C.byref(t)))
    return _tuning_dict(t)
