#!/usr/bin/env python3
"""Benchmark of the GPU join-order search (one JSON line on stdout).

Default workload (N=1): MPDP (DPsub with gpuqo_dpsub_bicc on) on the seeded
20-relation snowflake -- the headline configuration.  A step is one complete
plan of that query: one gq_plan call through the C ABI.  `value` is csg-cmp
pairs per second of the calls' wall time and `ms_per_step` their mean wall
time (host preparation, the query's upload, every kernel and the plan's
read-back included); the device time of the same calls (HIP events from the
first kernel to the plan walk) is reported beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Other workloads (for measurement, not the headline line): star12-dpsub,
clique18-dpsub, clique18-dpsize.
"""
import argparse
import ctypes as C
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "postgres-gpuqo_amd"))

METRIC = "MPDP plan ms + csg-cmp pairs/s (1 GPU); batched queries/s and p99 ms"
HBM_PEAK_GBS = 8000.0

# name: (generator, algorithm, bicc, CPU-baseline levels).  The CPU leg
# times whole plans, or -- where one plan takes the CPU minutes (the clique:
# about 30 min) -- the same query's DP through levels 2..L only.
WORKLOADS = {
    "mpdp-snowflake20": ("query_snowflake20", "dpsub", 1, None),
    "star12-dpsub": ("query_star12", "dpsub", 0, None),
    "clique18-dpsub": ("query_clique18", "dpsub", 0, 6),
    "clique18-dpsize": ("query_clique18", "dpsize", 0, 6),
}
# workloads whose parity is checked against the oracle's committed plans
# (the oracle needs about half an hour for one of them)
FIXTURE_LABELS = {"clique18-dpsub": "clique18-dpsub", "clique18-dpsize": "clique18-dpsize"}


TRAFFIC_PROFILES = {"mpdp-snowflake20": "k_eval_tree_mpdp-snowflake20.md"}


def measured_traffic(workload):
    """HBM bytes per launch of the dominant kernel from the committed rocprof
    summary (FETCH_SIZE and WRITE_SIZE runs scaled by the copy calibration),
    or None when no summary covers the workload."""
    name = TRAFFIC_PROFILES.get(workload)
    if not name:
        return None, None
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1]), "profiles/" + name
    except OSError:
        pass
    return None, None


def gather_roofline():
    """Measured random-access rate (GB/s of record bytes) per regime."""
    path = os.path.join(ROOT, "postgres-gpuqo_amd", "lib", "libgq_bench.so")
    L = C.CDLL(path)
    L.gq_gather_bench.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    out = {}
    for name, fp in (("l2_2MB", 2 << 20), ("mall_64MB", 64 << 20), ("hbm_4GB", 4 << 30)):
        for rec in (16, 40):
            g, ms = C.c_double(), C.c_double()
            rc = L.gq_gather_bench(fp, rec, 1 << 26, C.byref(g), C.byref(ms))
            if rc == 0:
                out[f"{rec}B@{name}"] = round(g.value, 1)
    return out


def batched_section(gpuqo_amd, abi, qgen, reps=5):
    """The 256-query batch: seeded mixed queries (n in 8..22; 40% snowflake, 30%
    star, 15% chain, 15% cycle), MPDP, planned as one batch (one launch per
    DP level bucket) vs one query at a time on the same GPU.  Both sides are
    timed around the C-ABI call (gq_plan_batch once / gq_plan per query),
    with the query descriptions built beforehand."""
    qs = qgen.queries_batch256()
    cfg = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
    args = gpuqo_amd.BatchArgs(qs)
    args.run(cfg)                                       # warm-up
    bt_wall, bt_dev = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        outs = args.run(cfg)
        bt_wall.append(time.perf_counter() - t0)
        bt_dev.append(outs[0].stats.gpu_ms / 1e3)
    batch_plans = [gpuqo_amd.plan_bits(gpuqo_amd.plan_tuples(outs[i])) for i in range(len(qs))]
    pairs = sum(outs[i].stats.pairs for i in range(len(qs)))
    for g in qs[:8]:
        gpuqo_amd.plan_raw(g, cfg)                      # warm-up
    lat, dev, same = [], [], True
    for i, g in enumerate(qs):
        t0 = time.perf_counter()
        out, _buf = gpuqo_amd.plan_raw(g, cfg)
        lat.append((time.perf_counter() - t0) * 1e3)
        dev.append(out.stats.gpu_ms)
        same = same and gpuqo_amd.plan_bits(gpuqo_amd.plan_tuples(out)) == batch_plans[i]
    lat_sorted = sorted(lat)
    bq = len(qs) / statistics.median(bt_wall)
    sq = len(qs) / (sum(lat) / 1e3)
    return {"workload": "batch256-mpdp", "queries": len(qs), "pairs": pairs,
            "batched_queries_per_s": bq,
            "batched_queries_per_s_device": len(qs) / statistics.median(bt_dev),
            "batch_ms_wall": statistics.median(bt_wall) * 1e3,
            "batch_p99_ms": sorted(bt_wall)[-1] * 1e3,
            "batched_pairs_per_s_device": pairs / statistics.median(bt_dev),
            "single_queries_per_s": sq,
            "single_queries_per_s_device": len(qs) / (sum(dev) / 1e3),
            "speedup_batched_vs_single": bq / sq,
            "single_p50_ms": lat_sorted[len(lat) // 2],
            "single_p99_ms": lat_sorted[min(len(lat) - 1, int(0.99 * len(lat)))],
            "batch_equals_single_all": same}


def serving_section(gpuqo_amd, abi, qgen, count=256, rate=1000.0, budget_gb=96, inflight=128, loads=(0.5, 0.8)):
    """The served mix: `count` seeded requests (70% snowflakes of 10-30 relations
    planned directly by MPDP, 30% snowflakes of 60-200 relations, depth 4,
    fanout 1-5, by UnionDP with k=25 and MPDP per partition) at the
    continuous-batching server.  Legs, after one untimed pass of the mix:
    capacity (every request submitted at once), Poisson arrivals at `rate`
    requests/s, and Poisson arrivals at each of `loads` x the measured
    capacity.  Latency is submission to completion (measured by the
    library); every served plan is compared with the same request planned
    alone by gq_plan, which also gives the one-at-a-time throughput (after
    one untimed pass as well)."""
    import random
    import threading
    mix = qgen.serving_mix(count, qgen.workload_seed("serve256"))
    direct = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
    union = abi.default_config(algorithm="dpsub", dpsub_bicc=1, idp_type="union", idp_n_iters=25)

    def poisson_leg(srv, lam, tag0):
        rng = random.Random(qgen.workload_seed("serve256") + int(lam))
        arrivals, t = [], 0.0
        for _ in mix:
            t += rng.expovariate(lam)
            arrivals.append(t)
        res, submitted = {}, [0]

        def collect():
            deadline = time.perf_counter() + 300.0
            while len(res) < count and time.perf_counter() < deadline:
                r = srv.poll(100)                  # returns at once while nothing is outstanding
                if r is None:
                    if submitted[0] == 0:
                        time.sleep(0.0002)
                    continue
                res[r[0] - tag0] = (r[1], gpuqo_amd.plan_bits(r[2]), r[3].wall_ms, time.perf_counter())

        th = threading.Thread(target=collect)
        th.start()
        t0 = time.perf_counter()
        for i, ((g, large), at) in enumerate(zip(mix, arrivals)):
            d = t0 + at - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(g, tag0 + i, union if large else direct)
            submitted[0] += 1
        th.join()
        lat = sorted(v[2] for v in res.values())
        span = max(v[3] for v in res.values()) - t0 if res else float("nan")
        q = lambda f: lat[min(len(lat) - 1, int(f * len(lat)))] if lat else float("nan")
        ok = len(res) == count and all(v[0] == 0 for v in res.values())
        return res, ok, {"arrival_rate_per_s": lam, "served_queries_per_s": count / span,
                         "latency_p50_ms": q(0.50), "latency_p99_ms": q(0.99),
                         "latency_max_ms": lat[-1] if lat else float("nan")}

    with gpuqo_amd.Server(direct, budget_gb << 30, max_inflight=inflight) as srv:
        # warm-up: the whole mix once, untimed (steady-state serving: the
        # request threads' and the engine's buffers at their working sizes)
        for i, (g, large) in enumerate(mix):
            srv.submit(g, 10 * count + i, union if large else direct)
        for _ in mix:
            if srv.poll(120000) is None:
                break
        # capacity: every request submitted at once
        t0 = time.perf_counter()
        for i, (g, large) in enumerate(mix):
            srv.submit(g, i, union if large else direct)
        sat = 0
        while sat < count:
            r = srv.poll(120000)
            if r is None:
                break
            sat += 1
        sat_s = time.perf_counter() - t0
        capacity = count / sat_s if sat == count else float("nan")
        res, ok, fixed = poisson_leg(srv, rate, 2 * count)
        loaded, loaded_res = [], []
        if capacity == capacity:
            for j, load in enumerate(loads):
                res2, ok2, leg = poisson_leg(srv, load * capacity, (4 + 2 * j) * count)
                loaded.append(dict(leg, load_of_capacity=load))
                loaded_res.append(res2)
                ok = ok and ok2
    # the same requests one at a time through gq_plan: throughput and parity
    # (after one untimed pass, as for the server)
    for g, large in mix:
        gpuqo_amd.plan_raw(g, union if large else direct)
    single_s, same = 0.0, ok
    for i, (g, large) in enumerate(mix):
        c = union if large else direct
        t1 = time.perf_counter()
        out, _buf = gpuqo_amd.plan_raw(g, c)
        single_s += time.perf_counter() - t1
        alone = gpuqo_amd.plan_bits(gpuqo_amd.plan_tuples(out))
        same = same and res.get(i, (None, None))[1] == alone
        for res2 in loaded_res:
            same = same and res2.get(i, (None, None))[1] == alone
    out = {"workload": "serve-mix256", "requests": count,
           "large_uniondp_requests": sum(1 for _, l in mix if l),
           "capacity_queries_per_s": capacity,
           "arrival_rate_per_s": rate, "served_queries_per_s": fixed["served_queries_per_s"],
           "latency_p50_ms": fixed["latency_p50_ms"], "latency_p99_ms": fixed["latency_p99_ms"],
           "latency_max_ms": fixed["latency_max_ms"],
           "loaded": loaded or None,
           "memo_budget_gb": budget_gb, "max_inflight": inflight, "all_ok": ok,
           "single_queries_per_s": count / single_s,
           "capacity_vs_single": capacity / (count / single_s),
           "served_equals_alone_all": same}
    return out


def host_info():
    """CPU model of this host and the oracle's compiler and flags."""
    import subprocess
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    flags = None
    try:
        with open(os.path.join(ROOT, "oracle", "Makefile")) as f:
            for line in f:
                if line.startswith("CXXFLAGS"):
                    flags = line.split("=", 1)[1].strip()
    except OSError:
        pass
    try:
        cxx = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        cxx = None
    return {"cpu_model": model, "compiler": cxx, "flags": flags}


class PlanCall:
    """Prebuilt arguments of one gq_plan call, so that a timed call is the
    C-ABI call itself."""

    def __init__(self, gpuqo_amd, abi, g, cfg):
        self.L = gpuqo_amd.lib()
        self.q = g.to_c()
        self.cfg = cfg
        self.nodes = (abi.PlanNode * (2 * g.n))()
        self.out = abi.PlanOut(C.cast(self.nodes, C.POINTER(abi.PlanNode)), 2 * g.n, 0)
        self.args = (C.byref(self.q), C.byref(self.cfg), C.byref(self.out))
        self.gpuqo_amd = gpuqo_amd

    def run(self, steps):
        """Wall ms and device ms of `steps` calls (host phases in self.host)."""
        wall, dev = [], []
        self.host = []
        for _ in range(steps):
            t1 = time.perf_counter()
            st = self.L.gq_plan(*self.args)
            wall.append((time.perf_counter() - t1) * 1e3)
            if st != 0:
                self.gpuqo_amd._check(st)
            s = self.out.stats
            dev.append(s.gpu_ms)
            self.host.append((s.host_prep_ms, s.host_enqueue_ms, s.host_post_ms))
        return wall, dev


def cpu_baseline(g, algo, bicc, levels=None, budget_s=12.0):
    """The oracle's CPU restatement timed on this host (bounded sample):
    whole plans, or with `levels` the DP through levels 2..levels."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # test infrastructure: CPU baseline leg only
    from gpuqo_amd import abi
    res = {}
    seq_algo = {("dpsub", 1): "cpu_dpsub_bicc", ("dpsub", 0): "cpu_dpsub", ("dpsize", 0): "cpu_dpsize"}[(algo, bicc)]

    def timed(cfg, budget):
        times, pairs = [], 0
        t_start = time.time()
        while True:
            if levels:
                st = oracle.dp_first_levels(g, cfg, levels)
                times.append(st.wall_ms / 1e3)
            else:
                t0 = time.perf_counter()
                _, _, st = oracle.plan(g, cfg)
                times.append(time.perf_counter() - t0)
            pairs = st.pairs
            if time.time() - t_start > budget or len(times) >= 20:
                break
        return statistics.median(times), len(times), pairs

    med, runs, pairs = timed(abi.default_config(algorithm=seq_algo), budget_s)
    res["seq"] = {"algorithm": seq_algo, "runs": runs, "median_ms": med * 1e3, "pairs": pairs,
                  "pairs_per_s": pairs / med, "levels": f"2..{levels}" if levels else "all"}
    if algo == "dpsub":
        # threads: the cores this process may run on, at most 16 (the host
        # share of one GPU)
        ncpu = max(1, min(len(os.sched_getaffinity(0)), 16))
        par_algo = "parallel_cpu_dpsub_bicc" if bicc else "parallel_cpu_dpsub"
        pcfg = abi.default_config(algorithm=par_algo, dpe_n_threads=ncpu, cpu_dpsub_parallel_chunk_size=256)
        pm, pruns, _ = timed(pcfg, budget_s / 2)
        res["parallel"] = {"algorithm": par_algo, "threads": ncpu, "chunk": 256, "runs": pruns,
                           "median_ms": pm * 1e3, "pairs_per_s": pairs / pm}
        if not levels:
            # dpe_dpccp: DPccp enumeration feeding the DPE worker pool
            # (gpuqo_dpe_pairs_per_depbuf 65536, run_all_generic.sh:63)
            dcfg = abi.default_config(algorithm="dpe_dpccp", dpe_n_threads=ncpu, dpe_pairs_per_depbuf=65536)
            dm, druns, _ = timed(dcfg, budget_s / 2)
            res["dpe_dpccp"] = {"algorithm": "dpe_dpccp", "threads": ncpu, "pairs_per_depbuf": 65536,
                                "runs": druns, "median_ms": dm * 1e3, "pairs_per_s": pairs / dm,
                                "note": "submits each pair twice in one orientation, as the reference's DPE"}
    return res


def fixture_plan(label):
    """The oracle's committed plan for a full-size workload."""
    with open(os.path.join(ROOT, "tests", "golden", "plans_large_v1.json")) as f:
        for r in json.load(f)["records"]:
            if r["label"] == label:
                return [(tuple(x[0]), x[1], x[2], x[3], x[4], x[5], x[6]) for x in r["plan"]]
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mpdp-snowflake20", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates byte counters) and exit")
    ap.add_argument("--batch-only", type=int, default=0, metavar="REPS",
                    help="only plan the 256-query batch (one warm-up, then REPS batches) and print its times")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch-shape knobs instead of tuning them on the workload first")
    ap.add_argument("--trees-only", type=int, default=0, metavar="REPS",
                    help="only plan the served mix's 40 direct snowflakes of 25-30 relations (rank memo) one at a "
                         "time, one warm-up pass and REPS timed passes, and print their times")
    ap.add_argument("--serve-only", action="store_true", help="only run the served section and print it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build plans on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    if args.copy_calibration:
        L = C.CDLL(os.path.join(ROOT, "postgres-gpuqo_amd", "lib", "libgq_bench.so"))
        L.gq_copy_bench.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        g, ms = C.c_double(), C.c_double()
        nbytes = args.copy_calibration << 20
        rc = L.gq_copy_bench(nbytes, 1, C.byref(g), C.byref(ms))
        print(json.dumps({"copy_bytes_each_way": nbytes, "launches": 2, "ms": ms.value, "gbps": g.value, "rc": rc}))
        return rc

    import gpuqo_amd
    from gpuqo_amd import abi, qgen
    if args.batch_only:
        qs = qgen.queries_batch256()
        bcfg = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
        bargs = gpuqo_amd.BatchArgs(qs)
        bargs.run(bcfg)
        wall, dev = [], []
        for _ in range(args.batch_only):
            t0 = time.perf_counter()
            outs = bargs.run(bcfg)
            wall.append((time.perf_counter() - t0) * 1e3)
            dev.append(outs[0].stats.gpu_ms)
        print(json.dumps({"workload": "batch256-mpdp", "queries": len(qs), "batch_ms_wall": wall,
                          "batch_ms_device": dev}), flush=True)
        return 0
    if args.trees_only:
        big = [g for g, large in qgen.serving_mix(256, qgen.workload_seed("serve256")) if not large and g.n >= 25]
        tcfg = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
        runs = []
        for rep in range(args.trees_only + 1):
            dev, pairs, kinds = 0.0, 0, set()
            for g in big:
                _, st = gpuqo_amd.plan(g, tcfg)
                dev += st.gpu_ms
                pairs += st.pairs
                kinds.add(st.memo_kind)
            if rep:
                runs.append(dev)
        print(json.dumps({"workload": "serve-direct-snowflakes25-30", "queries": len(big), "pairs": pairs,
                          "memo_kinds": sorted(kinds), "device_ms": runs,
                          "pairs_per_s_device": pairs / (statistics.median(runs) / 1e3)}), flush=True)
        return 0
    if args.serve_only:
        print(json.dumps(serving_section(gpuqo_amd, abi, qgen)), flush=True)
        return 0
    cfg_name, algo, bicc, cpu_levels = WORKLOADS[args.workload]
    g = getattr(qgen, cfg_name)()
    cfg = abi.default_config(algorithm=algo, dpsub_bicc=bicc)
    cfg_prof = abi.default_config(algorithm=algo, dpsub_bicc=bicc, profile_kernels=1)
    print(f"# workload {args.workload}: {g.shape}-{g.n}, algorithm {algo} bicc={bicc}", file=sys.stderr, flush=True)

    call = PlanCall(gpuqo_amd, abi, g, cfg)
    tuning = None
    if not args.no_autotune and g.shape in ("snowflake", "star"):
        # (the single-plan knobs shape the tree kernels only)
        # the default knobs first (reported beside the tuned number), then
        # gq_autotune on samples that are not the timed query: the
        # single-plan knobs on six other seeded queries of the same shape
        # and size, the batch knobs on the 256-query batch
        call.run(max(1, args.warmup))
        uw, ud = call.run(args.steps)
        t0 = time.perf_counter()
        with_batch = not args.no_batch and args.workload == "mpdp-snowflake20"
        sample = [getattr(qgen, g.shape)(g.n, qgen.workload_seed("tune") + i) for i in range(6)]
        knobs = gpuqo_amd.autotune(sample, cfg, reps=5, knobs=gpuqo_amd.TUNE_SINGLE)
        if with_batch:
            knobs = gpuqo_amd.autotune(qgen.queries_batch256(), abi.default_config(algorithm="dpsub", dpsub_bicc=1),
                                       reps=3, knobs=gpuqo_amd.TUNE_BATCH)
        tuning = {"knobs": knobs, "tune_s": round(time.perf_counter() - t0, 2),
                  "sample": {"single": f"6 seeded {g.shape}-{g.n} queries (not the timed one)",
                             "batch": "batch256-mpdp" if with_batch else None},
                  "untuned": {"ms_per_step": statistics.mean(uw), "plan_ms_wall_median": statistics.median(uw),
                              "plan_ms_device_median": statistics.median(ud)}}
        print(f"# autotune {knobs} in {tuning['tune_s']} s", file=sys.stderr, flush=True)
    call.run(args.warmup)
    # timed region: K gq_plan calls (no per-level events), each timed around
    # the C-ABI call; device time is the HIP-event span from the first kernel
    # to the plan walk on the DP stream
    t0 = time.perf_counter()
    wall_ms, dev_ms = call.run(args.steps)
    total_wall = time.perf_counter() - t0
    host_phases = {k: statistics.median(x[i] for x in call.host)
                   for i, k in enumerate(("prep_ms", "enqueue_ms", "post_ms"))}
    nodes, stats = gpuqo_amd.plan_tuples(call.out), call.out.stats
    print(f"# timed {args.steps} steps in {total_wall:.3f} s", file=sys.stderr, flush=True)
    # per-phase / per-launch times from separate profiled calls (HIP events
    # around every launch on the DP stream), outside the timed region
    prof_steps = max(5, min(args.steps, 20))
    f_ms, e_ms, f_n, e_n = 0.0, 0.0, 0, 0
    for _ in range(prof_steps):
        _, pst = gpuqo_amd.plan(g, cfg_prof)
        f_ms += pst.filter_ms
        e_ms += pst.eval_ms
        f_n += pst.filter_launches
        e_n += pst.eval_launches

    P, S, Cn = stats.pairs, stats.sets, stats.candidates
    wall_total_s = sum(wall_ms) / 1e3
    dev_total_s = sum(dev_ms) / 1e3
    value = P * args.steps / wall_total_s
    if algo == "dpsize":
        # DPsize: 72 B per candidate pair (two 36-B list entries), 128 B per
        # valid pair (two 64-B JoinRelationDpsize reads), 72 B per new set;
        # Cn counts the ordered candidate pairs the scan tests
        eval_bytes = 72.0 * Cn + 128.0 * P + 72.0 * S
        filt_bytes = 0.0
    else:
        # DPsub/MPDP (80 P + 40 S + 16 C): two memo probes of an 8-byte key
        # and a 32-byte JoinRelation per pair, one memo write per set, one
        # 8-byte candidate write + read per generated set (C: the 2^n - n - 1
        # integers the filter tests, or the connected sets that expansion
        # builds on tree graphs, which need no filter).
        eval_bytes = 80.0 * P + 40.0 * S
        filt_bytes = 16.0 * Cn
    eval_kernel = {"mpdp-snowflake20": "k_eval_tree", "star12-dpsub": "k_eval_tree",
                   "clique18-dpsub": "k_eval_wg<false> + k_eval_dpsub_chunk",
                   "clique18-dpsize": "k_ds_* (one level: k_ds_pairs and its six helpers)"}[args.workload]
    if e_ms >= f_ms:
        kern, kbytes, kms, kn = eval_kernel, eval_bytes, e_ms, e_n
    else:
        kern, kbytes, kms, kn = "k_filter_all", filt_bytes, f_ms, f_n
    per_launch_bytes = kbytes * prof_steps / kn
    avg_launch_ms = kms / kn
    achieved = per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(args.workload) if kern == eval_kernel else (None, None)
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
            "kernel": kern,
            "algorithmic_bytes_per_launch": per_launch_bytes, "avg_launch_ms": avg_launch_ms,
            "launches_per_step": kn // prof_steps, "profiled_steps": prof_steps,
            "phase_ms_per_step": {"filter": f_ms / prof_steps, "eval": e_ms / prof_steps}}
    if not args.no_roofline:
        ra = gather_roofline()
        roof["measured_random_access_gbps"] = ra
        whole = (eval_bytes + filt_bytes) / (dev_total_s / args.steps) / 1e9
        roof["whole_dp_algorithmic_gbps"] = round(whole, 2)
        if "40B@mall_64MB" in ra:
            roof["frac_of_measured_mall_40B"] = round(whole / ra["40B@mall_64MB"], 4)
        if "40B@hbm_4GB" in ra:
            roof["frac_of_measured_hbm_40B"] = round(whole / ra["40B@hbm_4GB"], 4)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "csg-cmp pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": statistics.mean(wall_ms),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "shape": g.shape, "relations": g.n, "algorithm": algo,
                   "gpuqo_dpsub_bicc": bicc, "pairs": P, "sets": S, "candidates": Cn,
                   "parallelism": "single-gpu"},
        "timing": "wall time of each gq_plan call (C ABI); device time beside it",
        "plan_ms_wall_median": statistics.median(wall_ms),
        "plan_ms_device_median": statistics.median(dev_ms),
        "pairs_per_s_device": P * args.steps / dev_total_s,
        "host_phases_median": host_phases,
        "plan_ms_wall_p99": sorted(wall_ms)[min(len(wall_ms) - 1, int(0.99 * len(wall_ms)))],
        "roofline": roof,
        "tuning": tuning,
    }
    if not args.no_batch and args.workload == "mpdp-snowflake20":
        out["batched"] = batched_section(gpuqo_amd, abi, qgen)
        out["served"] = serving_section(gpuqo_amd, abi, qgen)
    if not args.no_cpu:
        cb = cpu_baseline(g, algo, bicc, cpu_levels)
        seq = cb["seq"]
        what = (f"{seq['runs']} full plans of the same query" if not cpu_levels else
                f"{seq['runs']} runs of the same query's DP through levels 2..{cpu_levels} "
                f"({seq['pairs']} pairs; a whole plan takes the CPU about half an hour)")
        out["cpu_baseline"] = {"value": seq["pairs_per_s"], "unit": "csg-cmp pairs/s", "cores": 1, "kind": "port",
                               "sample": f"{what} with {seq['algorithm']} (median {seq['median_ms']:.1f} ms)",
                               "plan_ms": seq["median_ms"], "host": host_info(),
                               "nproc": os.cpu_count(), "cores_available": len(os.sched_getaffinity(0))}
        if "parallel" in cb:
            out["cpu_baseline"]["parallel"] = cb["parallel"]
        if "dpe_dpccp" in cb:
            out["cpu_baseline"]["dpe_dpccp"] = cb["dpe_dpccp"]
    # parity of the last GPU plan (outside the timed region): against the
    # oracle's committed plan for the full-size workloads, else the oracle
    if args.workload in FIXTURE_LABELS:
        want = fixture_plan(FIXTURE_LABELS[args.workload])
        out["parity_vs_oracle"] = ("bit-exact (committed oracle plan)"
                                   if want is not None and gpuqo_amd.plan_bits(nodes) == want else "MISMATCH")
    elif not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        st, want, _ = oracle.plan(g, abi.default_config(algorithm=algo, dpsub_bicc=bicc))
        out["parity_vs_oracle"] = "bit-exact" if gpuqo_amd.plan_bits(nodes) == gpuqo_amd.plan_bits(want) else "MISMATCH"
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
