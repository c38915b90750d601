#!/usr/bin/env python3
"""Benchmark of the GPU join-order search (one JSON line on stdout).

Default workload (N=1): MPDP (DPsub with gpuqo_dpsub_bicc on) on the seeded
20-relation snowflake -- the headline configuration.  A step is one complete
plan of that query on the GPU.  `value` is csg-cmp pairs per second of
device time (inputs resident in HBM: the query blob is uploaded before the
first timed kernel; the plan tree is read back after the last); the
PCIe-inclusive wall time per call is reported beside it.

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

WORKLOADS = {
    "mpdp-snowflake20": ("query_snowflake20", "dpsub", 1),
    "star12-dpsub": ("query_star12", "dpsub", 0),
    "clique18-dpsub": ("query_clique18", "dpsub", 0),
    "clique18-dpsize": ("query_clique18", "dpsize", 0),
}


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


def serving_section(gpuqo_amd, abi, qgen, count=256, rate=1000.0, budget_gb=96, inflight=128):
    """The served mix: `count` seeded requests (70% snowflakes of 10-30 relations
    planned directly by MPDP, 30% snowflakes of 60-200 relations, depth 4,
    fanout 1-5, by UnionDP with k=25 and MPDP per partition) arriving as a
    Poisson process of `rate` requests/s at the continuous-batching server.
    Latency is submission to completion (measured by the library); every
    served plan is compared with the same request planned alone by gq_plan,
    which also gives the one-at-a-time throughput.  Both the server and the
    one-at-a-time leg plan the mix once untimed first."""
    import random
    import threading
    mix = qgen.serving_mix(count, qgen.workload_seed("serve256"))
    direct = abi.default_config(algorithm="dpsub", dpsub_bicc=1)
    union = abi.default_config(algorithm="dpsub", dpsub_bicc=1, idp_type="union", idp_n_iters=25)
    rng = random.Random(qgen.workload_seed("serve256"))
    arrivals, t = [], 0.0
    for _ in mix:
        t += rng.expovariate(rate)
        arrivals.append(t)
    res = {}
    with gpuqo_amd.Server(direct, budget_gb << 30, max_inflight=inflight) as srv:
        # warm-up: the whole mix once, untimed (steady-state serving: the
        # request threads' and the engine's buffers at their working sizes)
        for i, (g, large) in enumerate(mix):
            srv.submit(g, count + 1 + i, union if large else direct)
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

        submitted = [0]

        def collect():
            deadline = time.perf_counter() + 300.0
            while len(res) < count and time.perf_counter() < deadline:
                r = srv.poll(100)                  # returns at once while nothing is outstanding
                if r is None:
                    if submitted[0] == 0:
                        time.sleep(0.0002)
                    continue
                res[r[0]] = (r[1], gpuqo_amd.plan_bits(r[2]), r[3].wall_ms, time.perf_counter())

        th = threading.Thread(target=collect)
        th.start()
        t0 = time.perf_counter()
        for i, ((g, large), at) in enumerate(zip(mix, arrivals)):
            d = t0 + at - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(g, i, union if large else direct)
            submitted[0] += 1
        th.join()
    ok = len(res) == count and all(v[0] == 0 for v in res.values())
    lat = sorted(v[2] for v in res.values())
    span = max(v[3] for v in res.values()) - t0 if res else float("nan")
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
        same = same and res.get(i, (None, None))[1] == gpuqo_amd.plan_bits(gpuqo_amd.plan_tuples(out))
    q = lambda f: lat[min(len(lat) - 1, int(f * len(lat)))] if lat else float("nan")
    return {"workload": "serve-mix256", "requests": count, "arrival_rate_per_s": rate,
            "large_uniondp_requests": sum(1 for _, l in mix if l),
            "capacity_queries_per_s": count / sat_s if sat == count else float("nan"),
            "served_queries_per_s": count / span, "latency_p50_ms": q(0.50), "latency_p99_ms": q(0.99),
            "latency_max_ms": lat[-1] if lat else float("nan"), "memo_budget_gb": budget_gb,
            "max_inflight": inflight, "all_ok": ok,
            "single_queries_per_s": count / single_s,
            "capacity_vs_single": (count / sat_s) / (count / single_s) if sat == count else float("nan"),
            "served_equals_alone_all": same}


def cpu_baseline(g, algo, bicc, budget_s=12.0):
    """The oracle's CPU restatement timed on this host (bounded sample)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # test infrastructure: CPU baseline leg only
    from gpuqo_amd import abi
    res = {}
    seq_algo = {("dpsub", 1): "cpu_dpsub_bicc", ("dpsub", 0): "cpu_dpsub", ("dpsize", 0): "cpu_dpsize"}[(algo, bicc)]
    cfg = abi.default_config(algorithm=seq_algo)
    times, pairs = [], 0
    t_start = time.time()
    while True:
        t0 = time.perf_counter()
        st, nodes, stats = oracle.plan(g, cfg)
        times.append(time.perf_counter() - t0)
        pairs = stats.pairs
        if time.time() - t_start > budget_s or len(times) >= 20:
            break
    med = statistics.median(times)
    res["seq"] = {"algorithm": seq_algo, "runs": len(times), "median_ms": med * 1e3, "pairs": pairs,
                  "pairs_per_s": pairs / med}
    if algo == "dpsub":
        # threads: the cores this process may run on, at most 16 (the host
        # share of one GPU)
        ncpu = max(1, min(len(os.sched_getaffinity(0)), 16))
        par_algo = "parallel_cpu_dpsub_bicc" if bicc else "parallel_cpu_dpsub"
        pcfg = abi.default_config(algorithm=par_algo, dpe_n_threads=ncpu, cpu_dpsub_parallel_chunk_size=256)
        ptimes = []
        t_start = time.time()
        while True:
            t0 = time.perf_counter()
            st, nodes, stats = oracle.plan(g, pcfg)
            ptimes.append(time.perf_counter() - t0)
            if time.time() - t_start > budget_s / 2 or len(ptimes) >= 20:
                break
        pm = statistics.median(ptimes)
        res["parallel"] = {"algorithm": par_algo, "threads": ncpu, "chunk": 256, "runs": len(ptimes),
                           "median_ms": pm * 1e3, "pairs_per_s": pairs / pm}
        # dpe_dpccp: DPccp enumeration feeding the DPE worker pool
        # (gpuqo_dpe_pairs_per_depbuf 65536, run_all_generic.sh:63)
        dcfg = abi.default_config(algorithm="dpe_dpccp", dpe_n_threads=ncpu, dpe_pairs_per_depbuf=65536)
        dtimes = []
        t_start = time.time()
        while True:
            t0 = time.perf_counter()
            st, nodes, stats = oracle.plan(g, dcfg)
            dtimes.append(time.perf_counter() - t0)
            if time.time() - t_start > budget_s / 2 or len(dtimes) >= 20:
                break
        dm = statistics.median(dtimes)
        res["dpe_dpccp"] = {"algorithm": "dpe_dpccp", "threads": ncpu, "pairs_per_depbuf": 65536,
                            "runs": len(dtimes), "median_ms": dm * 1e3, "pairs_per_s": pairs / dm,
                            "note": "submits each pair twice in one orientation, as the reference's DPE"}
    return res


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
    cfg_name, algo, bicc = WORKLOADS[args.workload]
    g = getattr(qgen, cfg_name)()
    cfg = abi.default_config(algorithm=algo, dpsub_bicc=bicc)
    cfg_prof = abi.default_config(algorithm=algo, dpsub_bicc=bicc, profile_kernels=1)
    print(f"# workload {args.workload}: {g.shape}-{g.n}, algorithm {algo} bicc={bicc}", file=sys.stderr, flush=True)

    tuning = None
    if not args.no_autotune:
        # gq_autotune before anything is timed: the single-plan knobs on the
        # workload itself, the batch knobs on the 256-query batch; the chosen
        # knobs hold for every section below
        t0 = time.perf_counter()
        with_batch = not args.no_batch and args.workload == "mpdp-snowflake20"
        knobs = gpuqo_amd.autotune([g], cfg, reps=5, knobs=gpuqo_amd.TUNE_SINGLE)
        if with_batch:
            knobs = gpuqo_amd.autotune(qgen.queries_batch256(), abi.default_config(algorithm="dpsub", dpsub_bicc=1),
                                       reps=3, knobs=gpuqo_amd.TUNE_BATCH)
        tuning = {"knobs": knobs, "tune_s": round(time.perf_counter() - t0, 2),
                  "sample": {"single": f"{g.shape}-{g.n}", "batch": "batch256-mpdp" if with_batch else None}}
        print(f"# autotune {knobs} in {tuning['tune_s']} s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        gpuqo_amd.plan(g, cfg)
    # timed region: plain calls (no per-level events); device time is the
    # HIP-event span from the first kernel to the plan walk on the DP stream
    dev_ms, wall_ms = [], []
    nodes = stats = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        nodes, stats = gpuqo_amd.plan(g, cfg)
        wall_ms.append((time.perf_counter() - t1) * 1e3)
        dev_ms.append(stats.gpu_ms)
    total_wall = time.perf_counter() - t0
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
    dev_total_s = sum(dev_ms) / 1e3
    value = P * args.steps / dev_total_s
    # algorithmic bytes (80 P + 40 S + 16 C): two memo probes of an 8-byte key
    # and a 32-byte JoinRelation per pair, one memo write per set, one 8-byte
    # candidate write + read per unranked set.
    eval_bytes = 80.0 * P + 40.0 * S
    filt_bytes = 16.0 * Cn
    eval_kernel = {"mpdp-snowflake20": "k_eval_tree", "star12-dpsub": "k_eval_tree",
                   "clique18-dpsub": "k_eval_wg<false> + k_eval_dpsub_chunk",
                   "clique18-dpsize": "k_ds_pairs"}[args.workload]
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
        "ms_per_step": statistics.mean(dev_ms),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "shape": g.shape, "relations": g.n, "algorithm": algo,
                   "gpuqo_dpsub_bicc": bicc, "pairs": P, "sets": S, "candidates": Cn,
                   "parallelism": "single-gpu"},
        "plan_ms_device_median": statistics.median(dev_ms),
        "plan_ms_wall_median": statistics.median(wall_ms),
        "plan_ms_wall_p99": sorted(wall_ms)[min(len(wall_ms) - 1, int(0.99 * len(wall_ms)))],
        "roofline": roof,
        "tuning": tuning,
    }
    if not args.no_batch and args.workload == "mpdp-snowflake20":
        out["batched"] = batched_section(gpuqo_amd, abi, qgen)
        out["served"] = serving_section(gpuqo_amd, abi, qgen)
    if not args.no_cpu:
        cb = cpu_baseline(g, algo, bicc)
        seq = cb["seq"]
        out["cpu_baseline"] = {"value": seq["pairs_per_s"], "unit": "csg-cmp pairs/s", "cores": 1, "kind": "port",
                               "sample": f"{seq['runs']} full plans of the same query with {seq['algorithm']} "
                                         f"(median {seq['median_ms']:.1f} ms)",
                               "plan_ms": seq["median_ms"], "host": platform.processor() or platform.machine(),
                               "nproc": os.cpu_count(), "cores_available": len(os.sched_getaffinity(0))}
        if "parallel" in cb:
            out["cpu_baseline"]["parallel"] = cb["parallel"]
        if "dpe_dpccp" in cb:
            out["cpu_baseline"]["dpe_dpccp"] = cb["dpe_dpccp"]
        # parity of the last GPU plan against the oracle (outside the timed region)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        st, want, _ = oracle.plan(g, abi.default_config(algorithm=algo, dpsub_bicc=bicc))
        out["parity_vs_oracle"] = "bit-exact" if gpuqo_amd.plan_bits(nodes) == gpuqo_amd.plan_bits(want) else "MISMATCH"
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
