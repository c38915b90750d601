"""Memo snapshots (gq_plan_snapshot, gq_snapshot_*): the device-resident
copy of a finished DP's memo, in the caller's numbering, from which the
plan of any connected sub-query is walked without planning again.

GPU: for dense (tree and general graphs), rank and open-addressing memos
and the wide DP (a tree and a graph with cycles), the snapshot's plan is
gq_plan's bit for bit and it holds every connected set; the plan walked
for connected subsets equals the full plan's subtree where that subtree
is the set's entry, and its cost equals the oracle's optimum for the
sub-query induced by the subset, planned alone; export -> import gives the
same walks; sets that are not connected are refused.  CPU: malformed
exports are refused before any device work."""
import ctypes as C
import random
import struct

import numpy as np
import pytest

import gpuqo_amd
import oracle
from gpuqo_amd import abi, qgen
from helpers import assert_same_plan


def finite_cycle(n, seed):
    r = random.Random(seed)
    g = qgen.QueryGraph(n, "cycle")
    for i in range(n):
        g.rows[i] = float(r.randint(20, 200))
    for i in range(n):
        g.add_generic(i, (i + 1) % n, qgen.f32(r.uniform(1e-3, 1e-2)))
    return g.build()


def induced(g, S):
    """The sub-query of the relations in S (renumbered in increasing
    order): their base relations as they are, the classes with both members
    in S with their selectivities, FK targets and statistics."""
    rels = [i for i in range(g.n) if (S >> i) & 1]
    idx = {r: j for j, r in enumerate(rels)}
    sub = qgen.QueryGraph(len(rels), "induced")
    sub.base = g.base[rels].copy()
    et = np.zeros((len(rels), 1), dtype=np.uint64)
    offs, members, sels, fks, vars_ = [0], [], [], [], []
    for e in range(g.n_ec):
        ms = [int(x) for x in g.ec_members[g.ec_offsets[e]:g.ec_offsets[e + 1]]]
        if not all((S >> m) & 1 for m in ms):
            continue
        a, b = idx[ms[0]], idx[ms[1]]
        et[a, 0] |= np.uint64(1 << b)
        et[b, 0] |= np.uint64(1 << a)
        for k, m in enumerate(ms):
            members.append(idx[m])
            f = int(g.ec_fks[g.ec_offsets[e] + k][0])
            fk = 0
            for t in range(g.n):
                if (f >> t) & 1:
                    fk |= 1 << idx[t]
            fks.append([fk])
            vars_.append(g.ec_vars[g.ec_offsets[e] + k])
        offs.append(len(members))
        sels.append(g.ec_sels[e])
    sub.edge_table = et
    sub.n_words = 1
    sub.ec_offsets = np.array(offs, dtype=np.int32)
    sub.ec_members = np.array(members or [0], dtype=np.int32)
    sub.ec_sels = np.array(sels or [0.0], dtype=np.float32)
    sub.ec_fks = np.array(fks or [[0]], dtype=np.uint64).reshape(-1, 1)
    sub.ec_vars = np.array(vars_, dtype=g.ec_vars.dtype) if vars_ else g.ec_vars[:1].copy()
    sub.n_ec = len(sels)
    sub.edges = [(idx[int(g.ec_members[g.ec_offsets[e]])], idx[int(g.ec_members[g.ec_offsets[e] + 1])], "x", 0.0)
                 for e in range(g.n_ec)
                 if all((S >> int(m)) & 1 for m in g.ec_members[g.ec_offsets[e]:g.ec_offsets[e + 1]])]
    sub._c = None
    return sub


def random_connected(g, rng, size):
    adj = g.adjacency()
    S = 1 << rng.randrange(g.n)
    while bin(S).count("1") < size:
        nb = 0
        for v in range(g.n):
            if (S >> v) & 1:
                nb |= adj[v]
        nb &= ~S
        if not nb:
            break
        bits = [v for v in range(g.n) if (nb >> v) & 1]
        S |= 1 << rng.choice(bits)
    return S


CASES = [
    ("snowflake20-mpdp-dense", lambda: qgen.query_snowflake20(), dict(algorithm="dpsub", dpsub_bicc=1)),
    ("star12-dpsub-dense", lambda: qgen.query_star12(), dict(algorithm="dpsub", dpsub_bicc=0)),
    ("cycle16-dpsub-dense", lambda: finite_cycle(16, 16), dict(algorithm="dpsub", dpsub_bicc=0)),
    ("snowflake22-mpdp-rank", lambda: qgen.snowflake(22, 0x22), dict(algorithm="dpsub", dpsub_bicc=1,
                                                                     memo_kind="rank")),
    ("snowflake18-mpdp-hash", lambda: qgen.snowflake(18, 0x18), dict(algorithm="dpsub", dpsub_bicc=1,
                                                                     memo_kind="hash")),
    ("chain40-dpsub-wide", lambda: qgen.snowflake(40, 0x40, depth=39, fanout=1), dict(algorithm="dpsub",
                                                                                       dpsub_bicc=0)),
    ("cycle36-mpdp-wide", lambda: finite_cycle(36, 36), dict(algorithm="dpsub", dpsub_bicc=1)),
]


def _subtrees(nodes):
    """set -> (rows, startup, total, width) bits of every node of a plan"""
    return {rel[0]: (gpuqo_amd.fbits(a), gpuqo_amd.fbits(b), gpuqo_amd.fbits(c), w)
            for (rel, a, b, c, w, _l, _r) in nodes}


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,over", CASES, ids=[c[0] for c in CASES])
def test_gpu_snapshot_plans_and_subplans(name, make, over):
    g = make()
    cfg = abi.default_config(**over)
    alone, st = gpuqo_amd.plan(g, cfg)
    with gpuqo_amd.Snapshot(g, cfg) as snap:
        assert_same_plan(gpuqo_amd.plan_bits(snap.plan), gpuqo_amd.plan_bits(alone), name)
        n_rels, n_sets, dev_bytes = snap.info()
        assert n_rels == g.n and n_sets == st.sets + g.n and dev_bytes >= 32 * n_sets
        full = snap.subplan((1 << g.n) - 1)
        assert_same_plan(gpuqo_amd.plan_bits(full), gpuqo_amd.plan_bits(alone), name + " full")
        # every node of the full plan is the memo entry of its set
        sub_of_plan = _subtrees(alone)
        for S, vals in list(sub_of_plan.items())[:12]:
            got = snap.subplan(S)
            assert _subtrees(got)[S] == vals
        # random connected subsets: the optimum of the induced sub-query
        rng = random.Random(0x5a + g.n)
        for _ in range(6):
            S = random_connected(g, rng, rng.randint(2, min(g.n, 12)))
            got = snap.subplan(S)
            k = bin(S).count("1")
            assert len(got) == 2 * k - 1 and got[0][0][0] == S
            sub = induced(g, S)
            ost, want, _ = oracle.plan(sub, abi.default_config(algorithm="dpsub", dpsub_bicc=1))
            assert ost == 0
            assert gpuqo_amd.fbits(got[0][3]) == gpuqo_amd.fbits(want[0][3]), (name, hex(S))
        # export -> import: same walks
        data = snap.export()
        assert data[:8] == b"GQMEMO1\0"
        version, nr, ns, rb = struct.unpack_from("<IIQI", data, 8)
        assert (version, nr, ns, rb) == (1, g.n, n_sets, 32) and len(data) == 64 + 32 * n_sets
        sets = [struct.unpack_from("<Q", data, 64 + 32 * i)[0] for i in range(ns)]
        assert sets == sorted(sets, key=lambda x: (bin(x).count("1"), x)) and len(set(sets)) == ns
        with gpuqo_amd.Snapshot.load(data) as again:
            assert again.info()[:2] == (g.n, n_sets)
            for S in list(sub_of_plan)[:8]:
                assert gpuqo_amd.plan_bits(again.subplan(S)) == gpuqo_amd.plan_bits(snap.subplan(S))
            assert again.export() == data


@pytest.mark.gpu
def test_gpu_snapshot_refusals():
    g = qgen.query_star12()
    cfg = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
    with gpuqo_amd.Snapshot(g, cfg) as snap:
        with pytest.raises(gpuqo_amd.GpuqoError) as e:          # two dimensions: not connected
            snap.subplan((1 << 3) | (1 << 5))
        assert e.value.status == abi.GQ_ERR_INVALID
        with pytest.raises(gpuqo_amd.GpuqoError) as e:          # outside the query
            snap.subplan(1 << 12)
        assert e.value.status == abi.GQ_ERR_INVALID
        assert len(snap.subplan(1 << 4)) == 1                   # a base relation
    for over in (dict(algorithm="dpsize"), dict(algorithm="dpsub", idp_type="idp2", idp_n_iters=6)):
        with pytest.raises(gpuqo_amd.GpuqoError) as e:
            gpuqo_amd.Snapshot(qgen.snowflake(14, 3), abi.default_config(**over))
        assert e.value.status == abi.GQ_ERR_UNSUPPORTED
    nodes, _ = gpuqo_amd.plan(g, cfg)                            # the library stays usable
    assert len(nodes) == 23


def test_snapshot_import_refuses_malformed_exports():
    """Validation happens before any device work."""
    good = bytearray(64 + 32 * 3)
    good[0:8] = b"GQMEMO1\0"
    struct.pack_into("<IIQI", good, 8, 1, 2, 3, 32)
    struct.pack_into("<QQfffi", good, 64, 1, 0, 10.0, 0.0, 1.0, 8)
    struct.pack_into("<QQfffi", good, 96, 2, 0, 20.0, 0.0, 2.0, 8)
    struct.pack_into("<QQfffi", good, 128, 3, 1, 5.0, 0.0, 9.0, 16)
    bad = []
    b = bytearray(good); b[0] = ord("X"); bad.append(b)                       # magic
    b = bytearray(good); struct.pack_into("<I", b, 8, 2); bad.append(b)       # version
    bad.append(bytearray(good[:-1]))                                          # truncated
    b = bytearray(good); struct.pack_into("<Q", b, 128 + 8, 3); bad.append(b)  # left == set
    b = bytearray(good); struct.pack_into("<Q", b, 128, 7); bad.append(b)     # relation outside n = 2
    b = bytearray(good); struct.pack_into("<Q", b, 64 + 8, 2); bad.append(b)  # base relation with a left side
    L = gpuqo_amd.lib()
    for buf in bad:
        h = C.c_void_p()
        data = bytes(buf)
        assert L.gq_snapshot_import(data, len(data), C.byref(h)) == abi.GQ_ERR_INVALID
        assert not h.value
