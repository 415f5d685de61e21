"""Benchmark of the POA alignment path on one MI355X.

Workload (default): seeded windows of the convex per-call shape -- a 1.5 kb
uniform ACGT template, 30 reads of 1-2 kb at 15% mixed error, local
alignment with m=5 n=-4 g=-8 e=-6 q=-10 c=-4 -- every read aligned
(Align) and merged (AddAlignment) in order, then GenerateConsensus and the
MSA.  One step runs the whole batch of windows through the batched form
(one persistent launch per round of reads).  Alongside, on the same bytes:

* the plain per-call drop-in (AlignmentEngine::Align, one read at a time)
  on one window,
* the oracle (the repository's CPU restatement of the reference) on the
  host cores on a bounded sample of windows, as the CPU baseline.

Before anything else a child process allocates and fills 144 GB of device
memory and exits (--prewarm-gb 0 skips it): the first process to use a
fresh machine's device memory runs its kernels about 5% slower
(device_prewarm).

Prints one JSON line.  Inputs are resident on the host when a step starts
and every step ends with all device work complete (the window driver
synchronises its stream), so the timed region brackets whole-job work.

Roofline: each pass of the batched form launches one persistent kernel per
score width; the int16 launch runs on the compute stream (the two halves of
the windows take turns on it) and the int32 launch, when a pass has one,
beside it on a side stream.  Every launch is timed by HIP events recorded
around it on its own stream.  The roofline is that of the dominant kernel
(the score width with the most launch time): achieved = its algorithmic DP
bytes per launch / its average launch duration.  The HBM traffic per launch
comes from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the
same workload (profiles/traffic.json), when one matches.
"""
import argparse
import concurrent.futures
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import spoa_amd  # noqa: E402

PEAK_HBM_GBS = 8000.0


def measured_traffic(workload, n_windows, kernel):
    """HBM bytes per alignment launch measured by rocprofv3 for this
    workload and kernel (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("windows_per_step") != n_windows or rec.get("kernel") != kernel:
        return None
    return rec.get("traffic_bytes_per_launch")


def dominant_kernel_roofline(ctx, stats, params, windows):
    """Roofline of the dominant alignment kernel: the score width with the
    most launch time over the timed steps; its algorithmic DP bytes per
    launch over its average launch duration (HIP events per launch)."""
    best = None
    for w in (16, 32):
        ms = sum(s["kernel_ms%d" % w] for s in stats)
        if best is None or ms > best[1]:
            best = (w, ms)
    w, ms = best
    byts = sum(s["dp_bytes%d" % w] for s in stats)
    n = sum(s["launches%d" % w] for s in stats)
    achieved = byts / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    # the kernel of that width: the reference's int32 local convex jobs run
    # on int16 rows with 8-column segments, other int32 jobs on int32 rows;
    # launch configuration of that width for the batch's longest read
    sub = spoa_amd.AlignmentEngine.create(*params).subtype
    narrow = w == 32 and sub == 2 and params[0] == spoa_amd.SW
    bits = 16 if w == 16 or narrow else 32
    k, waves, split = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
    spoa_amd.lib().spoa_hip_config_split(ctx.device, max(1, windows.max_read_len()), w, 1, ctypes.byref(k),
                                         ctypes.byref(waves), ctypes.byref(split))
    # template arguments: subtype, stored type, columns a lane, workgroup
    # size, reference lanes, split-capable, scalar-engine rules
    name = "PoaAlignKernel<%d, %s, %d, %d, %d, %s, false>" % (
        sub, "short" if bits == 16 else "int", k.value, 64 if waves.value == 1 else 512, 16 if w == 16 else 8,
        "true" if split.value else "false")
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "kernel": name, "score_bits": w, "launches": n,
            "avg_launch_ms": ms / max(1, n), "bytes_per_launch": byts / max(1, n)}


def cpu_baseline(windows, params, strand, budget_s, threads):
    """Times a CPU engine on the host cores over as many windows as fit
    about budget_s seconds (at least one per thread): the reference itself
    (its AVX2 engine, compiled from its sources into oracle/_ref) where that
    library was built, else the oracle (the repository's scalar C++
    restatement).  Returns a cpu_baseline record, or None where neither
    library is present."""
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libspoa_ref.so")
    port_so = os.path.join(ROOT, "oracle", "liboracle.so")
    if os.path.exists(ref_so):
        kind, what = "reference", "the reference (spoa v4.0.8 SimdAlignmentEngine, AVX2 build)"
        fn = ctypes.CDLL(ref_so).ref_run_window
    elif os.path.exists(port_so):
        kind, what = "port", "the oracle (scalar C++ restatement of the reference)"
        fn = ctypes.CDLL(port_so).oracle_run_window
    else:
        return None
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int] * 7 + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int,
                                        ctypes.c_char_p, ctypes.c_uint32]
    full = tuple(params) + (None,) * (7 - len(params))
    t, m, n, g, e, q, c = full
    e = g if e is None else e
    if q is None:
        q, c = g, e

    def one(w):
        reads = windows.reads(w)
        data = "".join(reads).encode()
        offs = [0]
        for r in reads:
            offs.append(offs[-1] + len(r))
        arr = (ctypes.c_uint64 * len(offs))(*offs)
        out = ctypes.create_string_buffer(1 << 20)
        k = fn(t, m, n, g, e, q, c, len(reads), arr, data, int(strand), out, 1 << 20)
        return out.raw[:k].decode() if k >= 0 else None

    # one window per thread first, to size the sample to the budget
    t0 = time.perf_counter()
    done = 0
    results = {}
    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        nxt = 0
        futs = {}
        while nxt < windows.n_windows and len(futs) < threads:
            futs[ex.submit(one, nxt)] = nxt
            nxt += 1
        while futs:
            fin, _ = concurrent.futures.wait(list(futs), return_when=concurrent.futures.FIRST_COMPLETED)
            for f in fin:
                results[futs.pop(f)] = f.result()
                done += 1
            if time.perf_counter() - t0 < budget_s and nxt < windows.n_windows:
                while nxt < windows.n_windows and len(futs) < threads:
                    futs[ex.submit(one, nxt)] = nxt
                    nxt += 1
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "windows/s", "cores": threads, "kind": kind,
            "sample": "%d windows of the benchmark workload (Align + AddAlignment per read, consensus, MSA) "
                      "through %s, one window per thread, %d host threads, %.1f s" % (done, what, threads, el),
            "consensus": results}


def copy_probe(nbytes):
    """One plain device-to-device copy of nbytes through the HIP runtime the
    library itself loads (its counters calibrate HBM byte counts).  No other
    GPU runtime may share the process: torch, for one, ships a HIP runtime
    of its own under the same soname, and whichever loads first serves
    both."""
    spoa_amd.lib()
    rt = ctypes.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if rt.hipMalloc(ctypes.byref(src), nbytes) != 0 or rt.hipMalloc(ctypes.byref(dst), nbytes) != 0:
        raise RuntimeError("copy probe: hipMalloc failed")
    try:
        if rt.hipMemset(src, 1, nbytes) != 0 or rt.hipMemcpy(dst, src, nbytes, 3) != 0:  # 3: device to device
            raise RuntimeError("copy probe: copy failed")
        if rt.hipDeviceSynchronize() != 0:
            raise RuntimeError("copy probe: synchronize failed")
    finally:
        rt.hipFree(src)
        rt.hipFree(dst)


def pct(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(p / 100.0 * (len(xs) - 1)))))
    return xs[k]


def cpu_serving(jobs, arrivals, threads, budget_s):
    """The same arrival trace into a pool of host threads running a CPU
    engine (the reference's AVX2 build where present, else the oracle), one
    job per thread at a time.  Arrivals stop after budget_s and the pool is
    given another budget_s to finish; jobs still running then are counted
    as unfinished (their latency exceeds the reported p99).  Returns
    throughput and latency percentiles over the finished jobs."""
    import queue
    import threading
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libspoa_ref.so")
    port_so = os.path.join(ROOT, "oracle", "liboracle.so")
    if os.path.exists(ref_so):
        kind, fn = "reference", ctypes.CDLL(ref_so).ref_run_window
    elif os.path.exists(port_so):
        kind, fn = "port", ctypes.CDLL(port_so).oracle_run_window
    else:
        return None
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int] * 7 + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int,
                                        ctypes.c_char_p, ctypes.c_uint32]
    lat, lock = [], threading.Lock()
    todo = queue.Queue()
    t0 = time.perf_counter()

    def worker():
        while True:
            i = todo.get()
            if i is None:
                return
            reads, params = jobs[i]
            t, m, n, g, e, q, c = tuple(params) + (None,) * (7 - len(params))
            e = g if e is None else e
            if q is None:
                q, c = g, e
            data = "".join(reads).encode()
            offs = [0]
            for r in reads:
                offs.append(offs[-1] + len(r))
            arr = (ctypes.c_uint64 * len(offs))(*offs)
            out = ctypes.create_string_buffer(1 << 16)
            fn(t, m, n, g, e, q, c, len(reads), arr, data, 0, out, 1 << 16)
            with lock:
                lat.append(1000.0 * (time.perf_counter() - t0 - arrivals[i]))

    pool = [threading.Thread(target=worker, daemon=True) for _ in range(threads)]
    for th in pool:
        th.start()
    submitted = 0
    for i in range(len(jobs)):
        dt = arrivals[i] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        if time.perf_counter() - t0 > budget_s:
            break
        todo.put(i)
        submitted += 1
    for _ in pool:
        todo.put(None)
    deadline = time.perf_counter() + budget_s
    while time.perf_counter() < deadline:
        with lock:
            if len(lat) == submitted:
                break
        time.sleep(0.05)
    el = time.perf_counter() - t0
    with lock:
        done = list(lat)
    return {"kind": kind, "cores": threads, "submitted": submitted, "finished": len(done),
            "windows_per_s": len(done) / el, "p50_ms": pct(done, 50), "p99_ms": pct(done, 99),
            "sample": "the first %d jobs of the lowest offered load's trace, %.0f s of arrivals" % (submitted, budget_s)}


def serving_bench(args, threads):
    """Config 5: Poisson arrivals of serving-mix jobs at several offered
    loads; throughput and p50 / p99 latency (submit to consensus + MSA)."""
    cfg = spoa_amd.CONFIGS[5]
    rates = [float(x) for x in args.rates.split(",")]
    srv = spoa_amd.Server(threads=threads)
    sweep = []
    try:
        # warm-up: a few jobs through both parameter sets
        wj, wa = spoa_amd.serving_jobs(args.seed + 1, 8, 1000.0, template=(200, 2000), reads=(5, 20))
        spoa_amd.serve_trace(wj, wa, server=srv)
        for k, rate in enumerate(rates):
            n = max(8, int(rate * args.serve_seconds))
            jobs, arrivals = spoa_amd.serving_jobs(args.seed + 10 + k, n, rate)
            before = srv.stats()
            t0 = time.perf_counter()
            res, wall, st = spoa_amd.serve_trace(jobs, arrivals, server=srv)
            lat = [r.latency_ms for r in res]
            ok = sum(1 for r in res if r.status == 0)
            kern = st["kernel_ms"] - before["kernel_ms"]
            byts = st["dp_bytes"] - before["dp_bytes"]
            launches = st["launches"] - before["launches"]
            sweep.append({"offered_per_s": rate, "jobs": n, "completed": ok, "wall_s": wall,
                          "windows_per_s": ok / wall, "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                          "gcups_dp": (st["cells"] - before["cells"]) / (kern / 1e3) / 1e9 if kern else 0.0,
                          "kernel_ms": kern, "launches": launches,
                          "achieved_gbs": byts / (kern / 1e3) / 1e9 if kern else 0.0,
                          "deferred": st["deferred"] - before["deferred"],
                          "max_active": st["max_active_seen"]})
            print("serving rate %.1f/s: %d jobs, %.1f windows/s, p50 %.0f ms, p99 %.0f ms" % (
                rate, n, ok / wall, pct(lat, 50), pct(lat, 99)), file=sys.stderr, flush=True)
    finally:
        srv.close()
    top = sweep[-1]
    cpu = None
    if not args.no_cpu:
        jobs, arrivals = spoa_amd.serving_jobs(args.seed + 10, max(8, int(rates[0] * args.serve_seconds)), rates[0])
        cpu = cpu_serving(jobs, arrivals, threads, args.cpu_budget)
    achieved = top["achieved_gbs"]
    out = {
        "metric": "windows/s (serving: continuous batching, consensus + MSA per job)",
        "value": top["windows_per_s"], "unit": "windows/s", "n_gpus": 1, "steps": len(rates), "warmup": 1,
        "ms_per_step": 1000.0 * top["wall_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int16/int32", "data": "synthetic (seeded serving mix, Poisson arrivals)",
        "config": {"workload": cfg["name"], "rates_per_s": rates, "seconds_per_rate": args.serve_seconds,
                   "seed": args.seed, "host_threads": threads, "device_prewarm_gb": PREWARMED_GB},
        "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel": "PoaAlignKernel (serving rounds)",
                     "launches": top["launches"]},
        "sweep": sweep,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def percall_threads_child(args, threads):
    """The per-call drop-in from `threads` host threads, one engine each
    (the reference's threading model), over the first windows of the
    workload; prints one JSON record with every window's consensus."""
    cfg = spoa_amd.CONFIGS[args.config]
    n_windows = args.windows or {1: 1024, 2: 2048, 3: 4096, 4: 256}[args.config]
    windows = spoa_amd.config_windows(args.config, args.seed, n_windows)
    some = windows.subset(list(range(args.percall_threads_child)))
    t0 = time.perf_counter()
    pm = spoa_amd.run_windows(some, *cfg["params"], mode="per_call", strand=cfg.get("strand", False), msa=True,
                              threads=threads)
    el = time.perf_counter() - t0
    print(json.dumps({"threads": threads, "windows": some.n_windows, "windows_per_s": some.n_windows / el,
                      "gcups": pm.stats["cells"] / (el * 1e9), "consensus": list(pm.consensus)}))


def percall_threads(args, nt, nw, res):
    """Runs percall_threads_child in a child process whose HIP runtime has
    as many hardware queues as threads (GPU_MAX_HW_QUEUES; HIP's default of
    4 would serialise the engines' streams beyond four), as a multi-threaded
    caller of the drop-in deploys it."""
    env = dict(os.environ, GPU_MAX_HW_QUEUES=str(max(4, nt)))
    cmd = [sys.executable, os.path.abspath(__file__), "--config", str(args.config), "--seed", str(args.seed),
           "--threads", str(nt), "--percall-threads-child", str(nw)]
    if args.windows:
        cmd += ["--windows", str(args.windows)]
    out = subprocess.run(cmd, env=env, check=True, stdout=subprocess.PIPE, text=True).stdout
    rec = json.loads(out.strip().split("\n")[-1])
    cons = rec.pop("consensus")
    rec["hw_queues"] = max(4, nt)
    rec["same_consensus_as_batched"] = all(cons[w] == res.consensus[w] for w in range(len(cons)))
    return rec


_PREWARM = """
import sys, torch
free, total = torch.cuda.mem_get_info()
gb = min(int(sys.argv[1]), int(free * 0.6) >> 30)
bufs = []
for _ in range(gb // 8):
    b = torch.empty(8 << 30, dtype=torch.uint8, device="cuda:0")
    b.fill_(0)
    bufs.append(b)
torch.cuda.synchronize()
print(len(bufs) * 8, flush=True)
"""

PREWARMED_GB = 0


def device_prewarm(gb):
    """Device memory that no process has used since the machine started
    makes the whole first process that uses it run its kernels about 5%
    slower (the default workload's int16 launches: 17.9 ms in the first
    process on each of six fresh machines, 16.9-17.5 ms in any process after
    one that had allocated and filled 144 GB; a process that only streamed
    an 8 GB buffer for 8 s did not help).  A child process allocates and
    fills `gb` GB (at most 60% of the free memory) and exits before the
    benchmark starts, outside the timed region, as a service would do once
    when its machine starts."""
    global PREWARMED_GB
    if gb <= 0:
        return
    try:
        out = subprocess.run([sys.executable, "-c", _PREWARM, str(gb)], capture_output=True, text=True,
                             timeout=300)
        if out.returncode == 0:
            PREWARMED_GB = int(out.stdout.strip().split()[-1])
        else:
            print("bench.py: device prewarm failed: %s" % out.stderr.strip()[-300:], file=sys.stderr)
    except (OSError, subprocess.SubprocessError, ValueError) as ex:
        print("bench.py: device prewarm failed: %s" % ex, file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--rates", default="10,25,50", help="config 5: offered loads, jobs/s")
    ap.add_argument("--serve-seconds", type=float, default=8.0, help="config 5: arrival span per load")
    ap.add_argument("--windows", type=int, default=None, help="windows per step (default per config)")
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--threads", type=int, default=0, help="host threads (0: all, at most 16)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-call", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch configurations)")
    ap.add_argument("--tune-rounds", type=int, default=8)
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (JSON)")
    ap.add_argument("--merge", choices=["auto", "host", "device"], default="auto",
                    help="where window graphs live and are merged: host threads, or device-resident "
                         "(auto: host for the 1.5 kb local convex windows, whose 9k-node graphs sort faster "
                         "on 16 host threads beside the alignments; device for the others)")
    ap.add_argument("--graph", action="store_true",
                    help="launch every pass of the host-merged batched form as a HIP graph (updated in place "
                         "from pass to pass)")
    ap.add_argument("--percall-threads-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="first run a plain device copy of this many MiB (calibrates HBM byte counters)")
    ap.add_argument("--prewarm-gb", type=int, default=144,
                    help="before anything else, a child process allocates and fills this many GB of device "
                         "memory and exits (0: skip); see device_prewarm")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(windows would shard across GPUs without any exchange between them)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if not args.percall_threads_child:
        device_prewarm(args.prewarm_gb)
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)

    threads = args.threads or min(16, os.cpu_count() or 1)
    if args.percall_threads_child:
        percall_threads_child(args, threads)
        return
    if args.config == 5:
        serving_bench(args, threads)
        return
    cfg = spoa_amd.CONFIGS[args.config]
    n_windows = args.windows or {1: 1024, 2: 2048, 3: 4096, 4: 256}[args.config]
    params = cfg["params"]
    strand = cfg.get("strand", False)
    windows = spoa_amd.config_windows(args.config, args.seed, n_windows)

    merge = args.merge if args.merge != "auto" else ("host" if args.config == 2 else "device")
    ctx = spoa_amd.BatchContext(*params, strand=strand, msa=True, threads=threads, resident=merge == "device",
                                graph=args.graph)
    tuned = None
    if not args.no_tune:
        # launch configuration per score width from the autotuner (untimed)
        tuned = ctx.autotune(windows, rounds=args.tune_rounds, cache=args.tune_cache)

    def step():
        return ctx.run(windows)

    for _ in range(args.warmup):
        step()
    stats = []
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        res = step()
        step_ms.append(1000.0 * (time.perf_counter() - ts))
        stats.append(res.stats)
    el = time.perf_counter() - t0
    ms_per_step = 1000.0 * el / args.steps
    wps = n_windows * args.steps / el
    cells = sum(s["cells"] for s in stats)
    dp_bytes = sum(s["dp_bytes"] for s in stats)
    kernel_ms = sum(s["kernel_ms"] for s in stats)
    launches = sum(s["launches"] for s in stats)
    roof = dominant_kernel_roofline(ctx, stats, params, windows)
    # the per-call engines below get the device memory the batched form held
    ctx.close()

    # plain per-call drop-in on the first window (same bytes)
    per_call = None
    if not args.no_per_call:
        one = windows.subset([0])
        t1 = time.perf_counter()
        pc = spoa_amd.run_windows(one, *params, mode="per_call", strand=strand, msa=True)
        pel = time.perf_counter() - t1
        per_call = {"windows_per_s": 1.0 / pel, "ms_per_window": 1000.0 * pel,
                    "gcups": pc.stats["cells"] / (pel * 1e9),
                    "kernel_ms": pc.stats["kernel_ms"], "align_calls": pc.stats["align_calls"],
                    "same_consensus_as_batched": pc.consensus[0] == res.consensus[0]}
        # the drop-in as a multi-threaded caller of the reference runs it: one
        # engine per host thread (made inside the timed region), a few
        # windows per thread in order
        nt = min(threads, n_windows)
        per_thread = {1: 8, 2: 2, 3: 4, 4: 1}[args.config]
        nw = min(n_windows, nt * per_thread)
        per_call["threads"] = percall_threads(args, nt, nw, res)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(windows, params, strand, args.cpu_budget, threads)
        if cpu is not None:
            agree = all(res.consensus[w] == c for w, c in cpu.pop("consensus").items())
            cpu["same_consensus_as_gpu"] = agree

    last = stats[-1]
    out = {
        "metric": "windows/s (batched GPU Align + AddAlignment + consensus + MSA)",
        "value": wps,
        "unit": "windows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16/int32",
        "data": "synthetic (seeded uniform ACGT templates, mixed substitution/insertion/deletion errors)",
        "config": {"workload": cfg["name"], "windows_per_step": n_windows, "seed": args.seed,
                   "params": list(params), "strand": strand, "host_threads": threads,
                   "graphs": "device-resident" if merge == "device" else "host",
                   "hip_graph": bool(args.graph), "device_prewarm_gb": PREWARMED_GB},
        "gcups_dp": cells / (kernel_ms / 1e3) / 1e9 if kernel_ms > 0 else 0.0,
        "gcups_end_to_end": cells / el / 1e9,
        "roofline": dict(roof, traffic=measured_traffic(cfg["name"], n_windows, roof["kernel"]),
                         all_launches={"launches": launches, "avg_launch_ms": kernel_ms / max(1, launches),
                                       "bytes_per_launch": dp_bytes / max(1, launches)}),
        "step_ms": [round(x, 1) for x in step_ms],
        "breakdown_ms_per_step": {k: last[k] for k in ("wall_ms", "align_ms", "kernel_ms", "h2d_ms",
                                                       "d2h_ms", "pack_ms", "alloc_ms", "wait_ms",
                                                       "unpack_ms", "graph_ms", "finish_ms")},
        "autotuned": tuned,
        "per_call": per_call,
        "batched_vs_per_call": (wps / per_call["windows_per_s"]) if per_call else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
