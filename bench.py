"""Benchmark: earliest-arrival queries/s with p50/p99 latency on one MI355X,
batched (64 queries per wavefront) against the plain one-query-per-launch
GPU path, with the CPU oracle timed on the host cores as a baseline.

Default (workload 5, the 500k-stop serving configuration): queries of the
workload's 60/25/15 point/range/intermodal mix arrive as a Poisson stream at
--rate queries/s (1000 by default) and go through the continuous-batching
serving loop.  A step is one batch of --batch queries of that stream; the
K timed batches arrive back to back as one stream.  value = queries / wall
time of the stream, latency = arrival -> completion per query (p50/p99
overall and per query kind).  A second stream of --low-queries queries at
--low-rate queries/s (below the server's capacity) reports latency without
the queueing of the saturated stream (outside the timed region).  The plain
engine runs a sample of the same stream one query at a time.

Other workloads (--workload 1..4) keep the batch modes: a step is one batch
of `--batch` queries; point-query batches are admitted whole once the
previous batch has finished (their preparation overlapping its rounds),
mixes with range queries take free slots as they appear (`--continuous`);
latency is admission -> completion there, the wait for admission reported
separately.  `--batch-per-step` runs each step as its own serving call.
Inputs (timetable image, query state) are resident on the GPU before the
timed region starts.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    1: "2k-stop grid timetable (45x45), 200 lines, 20 trips/line, 7 days; point EA, max 7 transfers",
    2: "50k-stop random-geometric timetable, 5k lines x 100 trips, 365-day bitfields, "
       "footpath degree 4; point EA queries, max 14 transfers",
    3: "200k-stop timetable, 20k lines, 800k trips, degree 6; 70% point / 30% range (1-4 h), "
       "max 8 transfers",
    4: "200k-stop timetable; intermodal point queries, 1-50 start/dest offsets, max 8 transfers",
    5: "500k-stop timetable, 60k lines, 2.5M trips, degree 6; 60% point / 25% range / "
       "15% intermodal, max 8 transfers",
}


def sync():
    """Device-wide synchronisation through the HIP runtime the library links
    (a second runtime, such as torch's bundled one, must not be initialised
    next to it)."""
    import ctypes
    import nigiri_amd as na
    na.lib()
    if ctypes.CDLL("libamdhip64.so.7").hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def route_scan_bytes(work):
    """Algorithmic bytes of the route-scan kernel: 8 B per route stop the scan
    visits (nigiri's visited stops less those before a scan's first marked
    stop, which the kernel skips), 2 B per event examined, 8 B per route-day
    test, 12 B per transport-day test, 2 B per tmp write."""
    return (8 * (work["route_stops"] - work["skipped_stops"]) + 2 * work["events"]
            + 8 * work["route_day_tests"] + 12 * work["transport_day_tests"] + 2 * work["tmp_writes"])


def committed_traffic(workload, **cfg):
    """HBM bytes per route-scan launch from the committed rocprof PMC summary
    of this workload (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by a
    device copy of known size).  PMC counters need their own profiler runs, so
    the bench line carries the figure of the last profiled build, when the
    summary was taken on this configuration (scale, batch, steps, warm-up
    batches and groups); else null."""
    path = os.path.join(ROOT, "profiles", f"k_route_scan_workload{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    fields = {}
    for line in open(path):
        if ":" in line and not line.startswith(("#", "|", " ")):
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    try:
        for k, v in cfg.items():
            if float(fields.get(k, "nan")) != float(v):
                return None
        return float(fields["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(tt, qs, seconds):
    """The oracle (a C++ restatement of nigiri's CPU raptor, built -O3
    -march=native without byte counters) timed on this host's cores in the
    style of exe/benchmark.cc:210-270 (one state per thread, an atomic query
    counter), at nproc threads and at one thread, on a bounded sample of the
    workload's queries sized to about `seconds` of work each."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n_q = len(qs)
    nproc = os.cpu_count() or 1
    probe = qs.subset(list(range(min(8, n_q))))
    r = oracle.search_many(tt, probe, threads=1, baseline=True)
    per_q = max(1e-4, r["wall_ms"] / 1e3 / len(probe))
    out = {}
    for threads in (nproc, 1):
        n = int(min(n_q, max(threads * 2, seconds * threads / per_q)))
        r = oracle.search_many(tt, qs.subset(list(range(n))), threads=threads, baseline=True)
        out[threads] = (n, n / (r["wall_ms"] / 1e3), r)
    n, qps, r = out[nproc]
    n1, qps1, r1 = out[1]
    return {"value": round(qps, 2), "unit": "queries/s", "cores": nproc, "kind": "port",
            "sample": f"first {n} queries of the workload's stream, the oracle (C++ restatement of "
                      f"nigiri's raptor search incl. reconstruction, -O3 -march=native) on {nproc} "
                      f"threads: p50 {np.percentile(r['latency_ms'], 50):.1f} ms, p99 "
                      f"{np.percentile(r['latency_ms'], 99):.1f} ms per query",
            "one_thread": {"value": round(qps1, 2), "queries": n1,
                           "p50_ms": round(float(np.percentile(r1["latency_ms"], 50)), 2),
                           "p99_ms": round(float(np.percentile(r1["latency_ms"], 99)), 2)},
            "cpu_model": cpu_model()}


def with_traffic_rate(roof):
    """Adds the measured traffic's rate and fraction of peak to a roofline
    object (traffic: HBM bytes per launch, or None)."""
    t, ms = roof.get("traffic"), roof.get("avg_launch_ms")
    if t is not None and ms:
        roof["traffic_gbs"] = round(t / (ms * 1e-3) / 1e9, 1)
        roof["traffic_frac"] = round(roof["traffic_gbs"] / HBM_PEAK_GBS, 4)
    return roof


def knob_overrides(args):
    """--knobs name=value,...: launch-shape settings for the engine (grid_waves
    and the knobs of Engine.configure); results do not depend on them."""
    out = {}
    for kv in filter(None, args.knobs.split(",")):
        k, v = kv.split("=")
        out[k.strip()] = int(v)
    return out


def pct(a, p):
    return round(float(np.percentile(a, p)), 2) if len(a) else None


def served_main(args):
    """The metric on the serving workload (config 5 by default): queries of the
    workload's mix arrive as a Poisson stream at --rate per second; a step is
    one batch of --batch queries of that stream; the K timed batches arrive
    back to back as one stream through the continuous-batching loop.  value =
    queries completed / wall time of the stream; latency = arrival ->
    completion."""
    import nigiri_amd as na
    cfg = args.workload
    t0 = time.time()
    tt = na.Timetable.generate(cfg, scale=args.scale)
    B = args.batch
    n_total = (args.warmup + args.steps) * B
    # the queries of the main stream, then those of the below-capacity stream
    qs = na.Queries.generate(tt, cfg, n_total + args.low_queries)
    print(f"[bench] workload {cfg}: {tt.view.n_locations} locations, {tt.view.n_routes} routes, "
          f"{tt.view.n_transports} transports ({time.time() - t0:.1f}s incl. contraction hierarchy)",
          file=sys.stderr, flush=True)
    img = na.Image(tt)
    max_k = 10
    eng = na.Engine(img, lanes=64, groups=args.groups, max_k=max_k)
    eng.profile(True)
    eng.configure(admission_wait_ms=1.0, graphs=not args.eager, **knob_overrides(args))
    kinds = np.array(["intermodal" if qs[i].dest_match_mode == na.MATCH_INTERMODAL
                      else ("range" if qs[i].is_interval else "point")
                      for i in range(n_total + args.low_queries)])
    rng = np.random.default_rng(7)
    caps = dict(journey_capacity=16 * B * max(1, args.steps), leg_capacity=400 * B * max(1, args.steps),
                legs="raw")
    for w in range(args.warmup):
        arr = np.cumsum(rng.exponential(1000.0 / args.rate, B))
        eng.serve(qs.subset(list(range(w * B, (w + 1) * B))), arr - arr[0], **caps)
        print(f"[bench] warmup batch {w} done", file=sys.stderr, flush=True)
    warm_launches = eng.stats(reset=True)["route_scan_launches"]
    first = args.warmup * B
    idx = list(range(first, n_total))
    arr = np.cumsum(rng.exponential(1000.0 / args.rate, len(idx)))
    arr -= arr[0]
    sync()
    t_start = time.perf_counter()
    res, wall = eng.serve(qs.subset(idx), arr, **caps)
    sync()
    t_total = time.perf_counter() - t_start
    st = eng.stats(reset=True)
    n = len(idx)
    value = n / t_total
    lat = np.array([r["latency_ms"] for r in res])
    queue = np.array([r["queue_ms"] for r in res])
    kd = kinds[first:n_total]
    ok = sum(1 for r in res if r["status"] == 0)
    by_kind = {}
    for k in ("point", "range", "intermodal"):
        m = kd == k
        by_kind[k] = {"n": int(m.sum()), "p50_ms": pct(lat[m], 50), "p99_ms": pct(lat[m], 99),
                      "queue_p50_ms": pct(queue[m], 50)}
    print(f"[bench] {n} queries served in {t_total:.1f} s", file=sys.stderr, flush=True)

    # a second stream below the server's capacity: latency without queueing
    # for slots (not part of the timed region)
    low = None
    if args.low_queries > 0:
        lidx = list(range(n_total, n_total + args.low_queries))
        larr = np.cumsum(rng.exponential(1000.0 / args.low_rate, len(lidx)))
        larr -= larr[0]
        t_low = time.perf_counter()
        lres, _ = eng.serve(qs.subset(lidx), larr, **caps)
        t_low = time.perf_counter() - t_low
        llat = np.array([r["latency_ms"] for r in lres])
        lkd = kinds[n_total:]
        low = {"offered_qps": args.low_rate, "queries": len(lidx), "completed_qps": round(len(lidx) / t_low, 2),
               "p50_ms": pct(llat, 50), "p99_ms": pct(llat, 99),
               "by_kind": {k: {"n": int((lkd == k).sum()), "p50_ms": pct(llat[lkd == k], 50),
                               "p99_ms": pct(llat[lkd == k], 99)} for k in ("point", "range", "intermodal")}}
        print(f"[bench] below-capacity stream: {len(lidx)} queries at {args.low_rate}/s in {t_low:.1f} s",
              file=sys.stderr, flush=True)

    # a lone point query on an idle engine, batched and plain
    pts = [i for i in range(n_total) if kinds[i] == "point"][:5]
    plain = na.Engine(img, lanes=1, groups=1, max_k=max_k)
    plain.configure(graphs=not args.eager)
    plain.search(qs.subset([pts[0]]))  # warm-up
    lone_b, lone_p = [], []
    for i in pts:
        t = time.perf_counter()
        eng.search(qs.subset([i]))
        lone_b.append((time.perf_counter() - t) * 1e3)
        t = time.perf_counter()
        plain.search(qs.subset([i]))
        lone_p.append((time.perf_counter() - t) * 1e3)
    # batched route-scan dispatches in launch order (for matching a profiler's
    # per-dispatch records to the timed region)
    print(f"[bench] batched route-scan dispatches: warm-up {warm_launches}, timed "
          f"{st['route_scan_launches']}, after {eng.stats()['route_scan_launches']}",
          file=sys.stderr, flush=True)
    # the plain one-query-per-launch path on a sample of the same stream
    ps = qs.subset(idx[:args.plain_sample])
    sync()
    tp = time.perf_counter()
    plain.search(ps)
    sync()
    plain_qps = len(ps) / (time.perf_counter() - tp)
    plain.close()
    print(f"[bench] plain path: {plain_qps:.2f} q/s", file=sys.stderr, flush=True)

    launches = max(1, st["route_scan_launches"])
    alg_bytes = route_scan_bytes(st["work"])
    avg_ms = st["route_scan_ms"] / launches
    achieved = (alg_bytes / launches) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    cpu = None if args.no_cpu else cpu_baseline(tt, qs.subset(idx), args.cpu_seconds)
    line = {
        "metric": "earliest-arrival queries/s + p50/p99 latency, batched vs 1-query GPU, 1 MI355X",
        "value": round(value, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"config {cfg}: " + WORKLOADS[cfg], "scale": args.scale,
                   "mode": "served: Poisson arrivals through the continuous-batching loop",
                   "offered_qps": args.rate, "batch": B, "queries": n, "slots": 64 * args.groups,
                   "lanes": 64, "groups": args.groups, "max_k": max_k},
        "completed_ok": ok,
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
        "latency": "arrival -> completion (queue, preparation, lower bounds, rounds, legs)",
        "by_kind": by_kind,
        "below_capacity": low,
        "lone_point_query_ms": {"batched": round(float(np.median(lone_b)), 2),
                                "plain": round(float(np.median(lone_p)), 2)},
        "plain_1query_qps": round(plain_qps, 2),
        "plain_sample": len(ps),
        "speedup_vs_plain": round(value / plain_qps, 2) if plain_qps > 0 else None,
        "rounds": st["rounds"], "epochs": st["epochs"],
        "round_launch": "eager" if args.eager else "hipGraph",
        "lane_utilization": round((st["work"]["route_stops"] - st["work"]["skipped_stops"])
                                  / max(1, 64 * st["work"]["wave_stops"]), 4),
        "work": {k: int(v) for k, v in st["work"].items()},
        "roofline": with_traffic_rate({"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": committed_traffic(cfg, scale=args.scale, batch=B, steps=args.steps,
                                                  warmup=args.warmup, groups=args.groups),
                     "kernel": "k_route_scan",
                     "alg_bytes_per_launch": int(alg_bytes / launches),
                     "avg_launch_ms": round(avg_ms, 4), "launches": launches}),
        "transfer_kernel_avg_ms": round(st["transfer_ms"] / max(1, st["transfer_launches"]), 4),
        "cpu_baseline": cpu,
    }
    eng.close()
    img.close()
    print(json.dumps(line), flush=True)


def pong_main(args):
    """Range queries by pong (nigiri's pong_search: ping/pong pairs, the pong
    pruned by the ping's round times) on the plain engine, beside the range
    search of the same queries (plain and batched engine) and the oracle's
    pong and range search on one CPU thread; checks that both algorithms find
    the same journeys in each query's interval."""
    # HIP maps streams onto 4 hardware queues by default, which caps how many
    # engines' kernels overlap; read once when HIP initialises (below)
    if args.pong_threads > 4 and int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"
    import nigiri_amd as na
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cfg = args.workload
    tt = na.Timetable.generate(cfg, scale=args.scale)
    pool = na.Queries.generate(tt, cfg, 8 * args.pong)
    idx = [i for i in range(len(pool)) if pool[i].is_interval][:args.pong]
    qs = pool.subset(idx)
    n = len(qs)
    img = na.Image(tt)
    max_k = 16 if cfg in (1, 2) else 10
    plain = na.Engine(img, lanes=1, groups=2, max_k=max_k)
    plain.configure(graphs=not args.eager)  # the range search below replays graphs otherwise
    plain.pong(qs[0])  # warm-up
    print(f"[bench] pong: {n} range queries, warm-up done", file=sys.stderr, flush=True)

    def in_window(q, journeys):
        return sorted((min(j[:2]), max(j[:2]), j[3]) for j in journeys
                      if q.interval_from <= (min(j[:2]) if q.dir == na.DIR_FWD else max(j[:2])) < q.interval_to)

    pong_ms, pong_exec, pong_res = [], [], []
    for i in range(n):
        t = time.perf_counter()
        r = plain.pong(qs[i])
        pong_ms.append((time.perf_counter() - t) * 1e3)
        pong_exec.append(r["n_epochs"])
        pong_res.append(r["journeys"])
    print("[bench] pong on the plain engine done", file=sys.stderr, flush=True)
    concurrent = None
    if args.pong_threads > 1:
        # independent plain engines (own state and stream) driven from host
        # threads, as nigiri's benchmark runs several GPU states at once
        import threading
        engines = [na.Engine(img, lanes=1, groups=2, max_k=max_k) for _ in range(args.pong_threads)]
        for e in engines:
            e.pong(qs[0])
        got = [None] * n
        nxt = iter(range(n))
        lock = threading.Lock()

        def work(e):
            while True:
                with lock:
                    i = next(nxt, None)
                if i is None:
                    return
                got[i] = e.pong(qs[i])["journeys"]

        ths = [threading.Thread(target=work, args=(e,)) for e in engines]
        t = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        wall = time.perf_counter() - t
        for e in engines:
            e.close()
        concurrent = {"threads": args.pong_threads, "qps": round(n / wall, 2),
                      "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                      "results_equal_sequential": got == pong_res}
        print("[bench] concurrent pong done", file=sys.stderr, flush=True)
    range_ms, range_res = [], []
    for i in range(n):
        t = time.perf_counter()
        r = plain.search(qs.subset([i]))[0]
        range_ms.append((time.perf_counter() - t) * 1e3)
        range_res.append(r["journeys"])
    print("[bench] range search on the plain engine done", file=sys.stderr, flush=True)
    batched = na.Engine(img, lanes=64, groups=1, max_k=max_k)
    batched.configure(graphs=not args.eager)
    batched.search(qs.subset([0]))
    t = time.perf_counter()
    batched.search(qs)
    batched_ms = (time.perf_counter() - t) * 1e3
    # device objects are released here, not at interpreter exit (a profiler
    # may have finalised the HIP runtime by then)
    batched.close()
    plain.close()
    img.close()
    agree = sum(in_window(qs[i], pong_res[i]) == in_window(qs[i], range_res[i]) for i in range(n))
    cpu = {}
    if not args.no_cpu:
        m = min(n, args.pong_cpu)
        t = time.perf_counter()
        for i in range(m):
            oracle.pong(tt, qs[i])
        cpu["pong_qps"] = round(m / (time.perf_counter() - t), 2)
        t = time.perf_counter()
        for i in range(m):
            oracle.search(tt, qs[i])
        cpu["range_qps"] = round(m / (time.perf_counter() - t), 2)
        cpu.update(cores=1, kind="port", sample=f"first {m} of the {n} range queries, one thread")

    def stat(ms):
        a = np.array(ms)
        return {"qps": round(len(a) / (a.sum() / 1e3), 2), "p50_ms": round(float(np.percentile(a, 50)), 1),
                "p99_ms": round(float(np.percentile(a, 99)), 1)}

    print(json.dumps({
        "metric": "pong range queries/s (ping/pong searches, one query at a time), 1 MI355X",
        "workload": f"config {cfg}: " + WORKLOADS[cfg] + " (its range queries)", "scale": args.scale,
        "queries": n, "pong_plain_gpu": {**stat(pong_ms), "executes_per_query": round(float(np.mean(pong_exec)), 1)},
        "pong_plain_gpu_concurrent": concurrent,
        "range_plain_gpu": stat(range_ms),
        "range_batched_gpu": {"qps": round(n / (batched_ms / 1e3), 2), "batch": n},
        "cpu_oracle": cpu or None,
        "pong_equals_range_in_interval": f"{agree}/{n}",
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--batch", type=int, default=0,
                    help="queries per step (default: 512 served, 2048 in batch modes)")
    ap.add_argument("--rate", type=float, default=1000.0,
                    help="served mode: Poisson arrival rate, queries/s (config 5: 1k-50k)")
    ap.add_argument("--queries", type=int, default=10000)
    ap.add_argument("--plain-sample", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--admission-wait-ms", type=float, default=2000.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 512 MiB (scales rocprof FETCH/WRITE_SIZE)")
    ap.add_argument("--serve", type=float, default=0.0, metavar="RATE",
                    help="served measurement (the default for workload 5) at RATE queries/s "
                         "for any workload")
    ap.add_argument("--groups", type=int, default=32, help="64-query groups (serving mode)")
    ap.add_argument("--low-rate", type=float, default=150.0,
                    help="served mode: offered rate of the second, below-capacity stream (queries/s)")
    ap.add_argument("--low-queries", type=int, default=768,
                    help="served mode: queries of the below-capacity stream (0: none)")
    ap.add_argument("--continuous", action="store_true",
                    help="the timed steps' batches go through one serving run back to back: the "
                         "next batch's preparation and lower bounds overlap the current batch's "
                         "rounds (continuous batching); latency = admission -> completion")
    ap.add_argument("--batch-per-step", action="store_true",
                    help="each step is its own serving run of one batch (nothing overlaps "
                         "across steps); default: the timed steps' batches go through one "
                         "serving run, each batch admitted whole once the previous one has "
                         "finished, its preparation and lower bounds overlapping the previous "
                         "batch's rounds")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of autotuning them on the first batch")
    ap.add_argument("--pong", type=int, default=0, metavar="N",
                    help="measure pong range search on N range queries of the workload instead")
    ap.add_argument("--pong-threads", type=int, default=16, metavar="T",
                    help="--pong mode: also run the queries on T plain engines from T host threads")
    ap.add_argument("--pong-cpu", type=int, default=4, metavar="M",
                    help="range queries timed on the CPU oracle in --pong mode")
    ap.add_argument("--knobs", default="",
                    help="launch-shape settings name=value,... (grid_waves, route_claim, transfer_claim, "
                         "pack_max_entries, transfer_pack_max_entries, loose_lanes)")
    ap.add_argument("--eager", action="store_true",
                    help="launch each round's kernels one by one instead of replaying captured hipGraphs")
    args = ap.parse_args()
    # default admission: whole batches for point-query workloads (1, 2, 4);
    # free slots for mixes with range queries (3, 5), whose epochs run one
    # after another and would hold a whole batch back
    if not args.batch_per_step and not args.continuous and args.workload in (3, 5):
        args.continuous = True
    args.pipelined = not args.batch_per_step and not args.continuous
    if args.copy_calibration:
        import torch
        x = torch.ones(512 << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 512 << 20}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(queries shard across GPUs without communication; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    if args.serve > 0.0:
        args.rate = args.serve
    if args.pong == 0 and (args.serve > 0.0 or args.workload == 5):
        args.batch = args.batch or 512
        served_main(args)
        return
    args.batch = args.batch or 2048
    if args.pong > 0:
        pong_main(args)
        return

    import nigiri_amd as na
    cfg = args.workload
    t0 = time.time()
    tt = na.Timetable.generate(cfg, scale=args.scale)
    qs = na.Queries.generate(tt, cfg, args.queries)
    print(f"[bench] workload {cfg}: {tt.view.n_locations} locations, {tt.view.n_routes} routes, "
          f"{tt.view.n_transports} transports, {len(qs)} queries ({time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    img = na.Image(tt)
    max_k = 16 if cfg in (2,) else (9 if cfg == 1 else 10)
    groups = max(1, args.batch // 64)
    eng = na.Engine(img, lanes=64, groups=groups, max_k=max_k)
    eng.profile(True)
    # a step submits a whole batch at once: let preparation fill every slot
    # before the first epoch starts
    eng.configure(admission_wait_ms=args.admission_wait_ms, graphs=not args.eager,
                  admit_when_idle=args.pipelined, **knob_overrides(args))
    B = groups * 64
    n_q = len(qs)

    def batch(i):
        idx = [(i * B + j) % n_q for j in range(B)]
        return qs.subset(idx)

    zeros = np.zeros(B)
    tuned = None
    if not args.no_autotune:
        t_tune = time.time()
        chosen, trials = eng.autotune(batch(0))
        tuned = {"chosen": chosen, "seconds": round(time.time() - t_tune, 1),
                 "trials": [[k, v, ms] for k, v, ms in trials]}
        print(f"[bench] autotune: {chosen} ({tuned['seconds']} s)", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        eng.serve(batch(w), zeros, journey_capacity=8 * B, leg_capacity=320 * B, legs="raw")
        print(f"[bench] warmup step {w} done", file=sys.stderr, flush=True)
    eng.stats(reset=True)
    sync()
    lat = []
    n_journeys = 0
    serve_wall = 0.0
    # legs are written into the caller's arrays (sized for this workload:
    # under 8 journeys and 40 legs per query), not turned into Python objects
    caps = dict(journey_capacity=8 * B, leg_capacity=320 * B, legs="raw")
    queue = []
    t_start = time.perf_counter()
    if args.continuous or args.pipelined:
        # all K batches queued at the start; slots take the next query as they free
        idx = [((args.warmup + s) * B + j) % n_q for s in range(args.steps) for j in range(B)]
        allq = qs.subset(idx)
        caps_all = dict(journey_capacity=8 * B * args.steps, leg_capacity=320 * B * args.steps,
                        legs="raw")
        res, wall = eng.serve(allq, np.zeros(len(idx)), **caps_all)
        serve_wall += wall
        lat += [r["latency_ms"] - r["queue_ms"] for r in res]
        queue += [r["queue_ms"] for r in res]
        n_journeys += sum(len(r["journeys"]) for r in res)
        print(f"[bench] {args.steps} batches: {wall:.1f} ms", file=sys.stderr, flush=True)
    else:
        for s in range(args.steps):
            res, wall = eng.serve(batch(args.warmup + s), zeros, **caps)
            serve_wall += wall
            lat += [r["latency_ms"] for r in res]
            queue += [r["queue_ms"] for r in res]
            n_journeys += sum(len(r["journeys"]) for r in res)
            print(f"[bench] step {s}: {wall:.1f} ms", file=sys.stderr, flush=True)
    sync()
    t_total = time.perf_counter() - t_start
    st = eng.stats()
    qps = args.steps * B / t_total
    lat = np.array(lat)

    # plain one-query-per-launch GPU path on a sample of the same queries
    plain = na.Engine(img, lanes=1, groups=1, max_k=max_k)
    plain.configure(graphs=not args.eager)
    ps = qs.subset(list(range(min(args.plain_sample, n_q))))
    plain.search(ps.subset([0]))  # warm-up
    sync()
    tp = time.perf_counter()
    plain.search(ps)
    sync()
    plain_qps = len(ps) / (time.perf_counter() - tp)
    plain.close()

    # roofline of the dominant kernel (route scan), live from HIP events
    launches = max(1, st["route_scan_launches"])
    alg_bytes = route_scan_bytes(st["work"])
    avg_ms = st["route_scan_ms"] / launches
    achieved = (alg_bytes / launches) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0

    cpu = None if args.no_cpu else cpu_baseline(tt, qs, args.cpu_seconds)

    line = {
        "metric": "earliest-arrival queries/s + p50/p99 latency, batched vs 1-query GPU, 1 MI355X",
        "value": round(qps, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"config {cfg}: " + WORKLOADS[cfg], "scale": args.scale,
                   "mode": ("continuous" if args.continuous else
                            "pipelined batches" if args.pipelined else "batch per step"),
                   "batch": B, "lanes": 64, "groups": groups, "max_k": max_k,
                   "queries": n_q},
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "latency": ("admission -> completion (queries queued for a slot; queue wait separate)"
                    if args.continuous or args.pipelined
                    else "arrival -> completion (a batch arrives at step start)"),
        "queue_p50_ms": round(float(np.percentile(queue, 50)), 3),
        "queue_p99_ms": round(float(np.percentile(queue, 99)), 3),
        "plain_1query_qps": round(plain_qps, 2),
        "speedup_vs_plain": round(qps / plain_qps, 2) if plain_qps > 0 else None,
        "journeys": n_journeys,
        "lane_utilization": round((st["work"]["route_stops"] - st["work"]["skipped_stops"])
                                  / max(1, 64 * st["work"]["wave_stops"]), 4),
        "transfer_lane_utilization": round(st["work"]["transfer_stations"]
                                           / max(1, 64 * st["work"]["transfer_items"]), 4),
        "hazard_lanes_per_item": round(st["work"]["hazard_lanes"] / max(1, st["work"]["hazard_items"]), 4),
        "rounds": st["rounds"], "epochs": st["epochs"],
        "round_launch": "eager" if args.eager else "hipGraph",
        "launch_config": eng.get_config(), "autotune": tuned,
        "graph_rounds": int(st["graph_rounds"]), "graph_captures": int(st["graph_captures"]),
        "work": {k: int(v) for k, v in st["work"].items()},
        "host_ms_per_step": {"serve_call": round(serve_wall / args.steps, 1),
                             "journey_extraction": round(st["results_ms"] / args.steps, 1),
                             "reconstruction": round(st["reconstruct_ms"] / args.steps, 1),
                             "legs": round(st["legs_ms"] / args.steps, 1),
                             "reconstructed_journeys": int(st["reconstructed"])},
        "roofline": with_traffic_rate({"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": committed_traffic(cfg, scale=args.scale, batch=B, steps=args.steps,
                                                  warmup=args.warmup, groups=groups),
                     "kernel": "k_route_scan",
                     "alg_bytes_per_launch": int(alg_bytes / launches),
                     "avg_launch_ms": round(avg_ms, 4), "launches": launches}),
        "transfer_kernel_avg_ms": round(st["transfer_ms"] / max(1, st["transfer_launches"]), 4),
        "cpu_baseline": cpu,
    }
    eng.close()
    img.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
