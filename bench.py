#!/usr/bin/env python3
"""Benchmark of the enumeration path on one MI355X.

Workload (one "step"): the Chung-Lu graph chunglu1m (n = 1,134,890, gamma 2.2,
seed 2; 2,983,214 undirected edges, max degree 29,277) resident in HBM; the
4-clique query and then the diamond query, each one enumeration launch
sequence in sgm mode (counting, no tuple output -- the reference's sgm).

Reported:
  value        matches/s = (count_4clique + count_diamond) / device time
  roofline     the dominant kernel, lds_batch_kernel<1, *> (LDS batches of
               roots with at most 64 row entries; its two occupancy
               instantiations are one kernel here): the bytes its launches must read
               at least once (root records and rows, the row vertices' records
               and bitmap-build scan ranges; computed on the host per lane
               class) over its device time, measured with HIP events on the
               stream it runs on, against the 8 TB/s HBM peak.  "traffic" is
               its measured HBM bytes per step (committed rocprofv3
               FETCH_SIZE / WRITE_SIZE summary); "b_alg_per_step" is
               sum over roots of 4(d(u)+2) + sum over neighbours 4(d(v)+2)
               (peeled degrees), which charges every root with the whole row of
               every neighbour and is not what any kernel reads.
  cpu_baseline the oracle (CPU restatement of the reference enumeration,
               OpenMP over roots, count only, no summed levels) on the same
               prepared graphs, same two queries, with the host's CPU model,
               its logical CPU count and the threads used
  parity       GPU counts and tuple-multiset hashes (count-only and emitting
               runs) equal to the oracle's
  emit_hash    the same step with every sgm match walked and hashed

Other modes (each prints one JSON line of its own):
  --mode batched  rmat22cut: the 8 named central-node patterns (sgc), one
                  plain launch sequence per query vs one captured hipGraph with
                  autotuned group widths / donation thresholds
  --mode large    rmat26: 4-clique, W5 and 5-clique counts on the scale-26
                  R-MAT graph, one launch sequence each and then batched
  --mode serve    chunglu3m: a Poisson stream over the 207 central-node
                  patterns (20% sgm with a 1e6 tuple buffer) through the
                  serving loop vs the plain path, p50/p99 latency and queries/s
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=("plain", "batched", "serve", "large"), default="plain")
    ap.add_argument("--workload", default=None)
    ap.add_argument("--queries", type=int, default=160, help="serve: queries in the stream")
    ap.add_argument("--qps", type=float, default=0.0, help="serve: arrival rate (0: 0.8 x plain capacity)")
    ap.add_argument("--serve-lanes", type=int, default=4)
    ap.add_argument("--deadline-ms", type=float, default=1000.0, help="serve: per-query deadline")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--cold-sample", type=int, default=4, help="serve: queries timed the reference's way")
    ap.add_argument("--max-k", type=int, default=7, help="serve: only patterns with at most this many vertices")
    ap.add_argument("--tune-cap-ms", type=float, default=1000.0,
                    help="batched: queries slower than this alone keep the default launch parameters")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--large-queries", default="",
                    help="large: comma-separated subset of the workload's queries (default: all)")
    ap.add_argument("--large-batch-cap-ms", type=float, default=60000.0,
                    help="large: run the completed queries again as one batch only below this total")
    ap.add_argument("--once", action="store_true",
                    help="profiling mode: one launch sequence per query, no warmup, no CPU leg")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device (counter calibration)")
    return ap.parse_args()


def calib_copy(mib):
    """Device-to-device copy of a known size through the HIP runtime (the
    runtime's copy kernel), used to scale FETCH_SIZE / WRITE_SIZE readings."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, C.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"calibration copy: {n} bytes read, {n} bytes written", file=sys.stderr)


def read_traffic(workload):
    """HBM bytes from the committed FETCH_SIZE/WRITE_SIZE summary, if any:
    per step and per launch of the dominant kernel."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for ln in f:
            if "=" in ln and not ln.startswith("#"):
                k, v = ln.split("=", 1)
                try:
                    out[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    return out


def host_cpu():
    """CPU model, logical CPUs of the host and CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


# The kernel the roofline line is about: lds_batch_kernel<1, *>, the LDS batches
# of roots with at most 64 row entries (lane class 0, one DFS lane per item) -- the
# largest share of device time in the kernel trace of this workload.
DOMINANT_LG = 0
DOMINANT_KERNEL = "lds_batch_kernel<1, *> (narrowest lane class)"


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} requested; this build measures exactly one GPU "
              "(the path is single-GPU by design)",
              file=sys.stderr)
        sys.exit(2)
    if args.mode == "batched":
        return run_batched(args)
    if args.mode == "serve":
        return run_serve(args)
    if args.mode == "large":
        return run_large(args)
    args.workload = args.workload or "chunglu1m"
    import numpy as np

    import beep_amd as B
    from beep_amd import patterns as P
    from beep_amd import workloads as W

    wl = W.WORKLOADS[args.workload]
    t0 = time.time()
    g = W.make_graph(args.workload)
    t_gen = time.time() - t0
    queries = wl["queries"]
    preps = []
    t0 = time.time()
    for name, task in queries:
        s, d = P.directed(P.read_mtx_edges(name))
        plan = B.Plan.from_edges(s, d, B.mode_of(task))
        pp = B.Prepared(g, plan)
        preps.append((name, task, plan, pp))
    t_prep = time.time() - t0
    t0 = time.time()
    for _, _, _, pp in preps:
        pp.upload(0)
    t_upload = time.time() - t0
    descs = [pp.describe() for _, _, _, pp in preps]

    if args.once:
        for (name, task, plan, pp), dsc in zip(preps, descs):
            r = pp.count(emit=B.EMIT_NONE, lanes=args.lanes)
            print(json.dumps({"query": name, "count": r["count"], "t_enum_ms": r["t_enum_ms"],
                              "t_lds_class_ms": r["t_lds_class_ms"],
                              "lds_class_bytes": list(dsc.lds_class_bytes),
                              "alg_bytes": pp.alg_bytes, "donations": r["donations"]}))
        if args.calib_copy:
            calib_copy(args.calib_copy)
        return

    # correctness run with tuple hashes (untimed)
    gpu_ref = []
    for name, task, plan, pp in preps:
        r = pp.count(emit=B.EMIT_HASH, lanes=args.lanes)
        gpu_ref.append(r)

    def timed(emit):
        for _ in range(args.warmup):
            for _, _, _, pp in preps:
                pp.count(emit=emit, lanes=args.lanes)
        n = len(preps)
        tot, hub, lds, dom = [0.0] * n, [0.0] * n, [0.0] * n, [0.0] * n
        cnt, hs = [0] * n, [None] * n
        wall0 = time.time()
        for _ in range(args.steps):
            for i, (_, _, _, pp) in enumerate(preps):
                r = pp.count(emit=emit, lanes=args.lanes)
                tot[i] += r["t_enum_ms"]
                hub[i] += r["t_hub_ms"]
                lds[i] += r["t_lds_ms"]
                dom[i] += r["t_lds_class_ms"][DOMINANT_LG]
                cnt[i] = r["count"]
                hs[i] = (r["hash_sum"], r["hash_xor"])
        return dict(wall=time.time() - wall0, tot=tot, hub=hub, lds=lds, dom=dom, counts=cnt, hashes=hs)

    plain = timed(B.EMIT_NONE)
    emitting = timed(B.EMIT_HASH)
    K = args.steps
    # one profiled run per query (untimed): busy-time balance of workgroups,
    # waves and compute units
    balance = {}
    for (q, _), (_, _, _, pp) in zip(queries, preps):
        r = pp.count(emit=B.EMIT_NONE, lanes=args.lanes, profile=True)
        balance[q] = {"lds": round(r["balance_lds"], 3), "hub": round(r["balance_hub"], 3),
                      "cu": round(r["balance_cu"], 3), "cus_busy": r["cus_busy"],
                      "donated_frac": round(r["donated_frac"], 3),
                      "lds_build_frac": round(r["lds_build_frac"], 3)}
    dev_ms_step = sum(plain["tot"]) / K
    counts = plain["counts"]
    matches = sum(counts)
    value = matches / (dev_ms_step / 1000.0)
    b_alg = sum(pp.alg_bytes for _, _, _, pp in preps)
    # dominant kernel: compulsory bytes of its launches / their device time
    dom_bytes = sum(dsc.lds_class_bytes[DOMINANT_LG] for dsc in descs)
    dom_ms = sum(plain["dom"]) / K
    achieved = dom_bytes / (dom_ms / 1000.0) / 1e9 if dom_ms > 0 else None
    traffic = read_traffic(args.workload)
    dom_traffic = traffic.get("dominant_bytes_per_step")
    step_traffic = traffic.get("bytes_per_step")

    # CPU baseline: the oracle (the reference's enumeration restated, OpenMP over
    # roots, no summed levels) counting only, on the same prepared graphs; then,
    # untimed, the same with tuple hashes for parity with the GPU's.
    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        model, nproc, usable = host_cpu()
        cores = min(usable, 16)
        rp, col = g.csr()
        src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(rp).astype(np.int64))
        og = O.Graph(src, col)
        cpu_s = 0.0
        cpu_matches = 0
        parity = True
        for (name, task, plan, pp), gr in zip(preps, gpu_ref):
            s, d = P.directed(P.read_mtx_edges(name))
            op = O.Prepared(og, O.plan(s, d, B.mode_of(task)))
            t0 = time.time()
            o = op.count(nthreads=cores, want_hash=False)
            cpu_s += time.time() - t0
            cpu_matches += o["count"]
            oh = op.count(nthreads=cores, want_hash=True)
            ok = (o["count"] == gr["count"] == oh["count"] and oh["hash_sum"] == gr["hash_sum"]
                  and oh["hash_xor"] == gr["hash_xor"])
            parity = parity and ok
        cpu = {"value": cpu_matches / cpu_s, "unit": "matches/s", "cores": cores, "kind": "port",
               "nproc": nproc, "cpus_usable": usable, "cpu_model": model,
               "sample": f"full workload ({' then '.join(q for q, _ in queries)}) on the same "
                         f"prepared graphs, count only, {cpu_s:.2f} s of enumeration on {cores} "
                         f"threads (the host grants 16 CPUs per GPU)"}
        parity = (bool(parity) and counts == [r["count"] for r in gpu_ref]
                  and emitting["counts"] == counts
                  and emitting["hashes"] == [(r["hash_sum"], r["hash_xor"]) for r in gpu_ref])

    emit_ms = sum(emitting["tot"]) / K
    out = {
        "metric": "subgraph matches/s",
        "value": value,
        "unit": "matches/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": plain["wall"] * 1000.0 / K,
        "device_ms_per_step": dev_ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: Chung-Lu n=1,134,890, gamma 2.2, seed 2, "
                        f"{g.m // 2} undirected edges, max degree {g.max_degree}; "
                        + ", then ".join(f"{q} {t}" for q, t in queries) + ", one query per launch",
            "n": g.n, "directed_entries": g.m, "queries": [q for q, _ in queries],
            "counts": counts, "lanes": args.lanes or "auto",
        },
        "roofline": {"bound": "hbm", "kernel": DOMINANT_KERNEL,
                     "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS if achieved else None,
                     "traffic": dom_traffic / max(len(preps), 1) if dom_traffic else None,
                     "alg_bytes_per_launch": dom_bytes / max(len(preps), 1),
                     "launches_per_step": len(preps),
                     "kernel_ms_per_step": dom_ms,
                     "measured_gbs": dom_traffic / (dom_ms / 1000.0) / 1e9 if dom_traffic and dom_ms else None,
                     "step_traffic": step_traffic,
                     "step_measured_gbs": (step_traffic / (dev_ms_step / 1000.0) / 1e9
                                           if step_traffic else None),
                     "b_alg_per_step": b_alg},
        "cpu_baseline": cpu,
        "parity": parity,
        "emit_hash": {"value": sum(emitting["counts"]) / (emit_ms / 1000.0), "unit": "matches/s",
                      "device_ms_per_step": emit_ms,
                      "note": "sgm match sets: every tuple walked and hashed, no summed last level"},
        "per_query_ms": {q: plain["tot"][i] / K for i, (q, _) in enumerate(queries)},
        "per_query_phase_ms": {q: {"hub": plain["hub"][i] / K, "lds": plain["lds"][i] / K,
                                   "dominant_kernel": plain["dom"][i] / K}
                               for i, (q, _) in enumerate(queries)},
        "balance": balance,
        "host_s": {"generate": round(t_gen, 3), "prepare": round(t_prep, 3),
                   "upload": round(t_upload, 3)},
    }
    print(json.dumps(out))


def pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def run_batched(args):
    """Eight named patterns (sgc) on rmat22cut: plain launches vs one hipGraph."""
    import beep_amd as B
    from beep_amd import autotune as T
    from beep_amd import patterns as P
    from beep_amd import workloads as W

    name = args.workload or "rmat22cut"
    wl = W.WORKLOADS[name]
    t0 = time.time()
    g = W.make_graph(name)
    print(f"[bench] generated {name} in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    queries = wl["queries"]
    preps = []
    t0 = time.time()
    for q, task in queries:
        s, d = P.directed(P.read_mtx_edges(q))
        pp = B.Prepared(g, B.Plan.from_edges(s, d, B.mode_of(task)))
        pp.upload(0)
        preps.append(pp)
    t_prep = time.time() - t0
    print(f"[bench] prepared {len(preps)} queries in {t_prep:.1f} s", file=sys.stderr, flush=True)
    # plain: one launch sequence per query, one after the other
    plain = []
    for (q, _), pp in zip(queries, preps):
        pp.count()  # warm
        r = pp.count()
        plain.append(r)
        print(f"[bench] plain {q}: {r['t_enum_ms']:.2f} ms count {r['count']}", file=sys.stderr, flush=True)
    plain_ms = sum(r["t_enum_ms"] for r in plain)
    # autotune (group width x donation threshold) on the queries that are cheap to sweep
    tuned = []
    t0 = time.time()
    for (q, _), pp, r in zip(queries, preps, plain):
        if r["t_enum_ms"] > args.tune_cap_ms:
            tuned.append({"lanes": 0, "donate_min": 64, "build_mode": 0, "ms": r["t_enum_ms"], "tuned": False})
            continue
        best, ms, _ = T.tune(pp, reps=1)
        tuned.append(dict(best, ms=ms, tuned=True))
        print(f"[bench] tuned {q}: {best} {ms:.2f} ms", file=sys.stderr, flush=True)
    t_tune = time.time() - t0
    batch = B.Batch(preps, tune=tuned)
    res, t_first = batch.run()
    ts = []
    for _ in range(max(1, args.steps // 4)):
        res, t = batch.run()
        ts.append(t)
        print(f"[bench] batched run {t:.2f} ms", file=sys.stderr, flush=True)
    batched_ms = min(ts)
    counts = [r["count"] for r in res]
    parity = counts == [r["count"] for r in plain]
    golden = os.path.join(ROOT, "tests", "golden", f"{name}_counts.json")
    if os.path.exists(golden):
        with open(golden) as f:
            ref = json.load(f)
        parity = parity and all(ref.get(q, c) == c for (q, _), c in zip(queries, counts))
    out = {
        "metric": "batched queries/s",
        "value": len(queries) / (batched_ms / 1000.0),
        "unit": "queries/s",
        "n_gpus": 1, "steps": len(ts), "warmup": 1,
        "ms_per_step": batched_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"{name}: R-MAT scale 22, ids >= 3,774,768 rejected, seed 3, "
                               f"{g.m // 2} undirected edges; 8 central-node patterns, sgc",
                   "queries": [q for q, _ in queries], "counts": counts},
        "plain": {"value": len(queries) / (plain_ms / 1000.0), "unit": "queries/s", "device_ms": plain_ms,
                  "per_query_ms": {q: r["t_enum_ms"] for (q, _), r in zip(queries, plain)}},
        "batched_ms_first_launch": t_first,
        "speedup_vs_plain": plain_ms / batched_ms,
        "autotune": {"params": {q: t for (q, _), t in zip(queries, tuned)}, "seconds": round(t_tune, 2),
                     "space": {"lanes": list(T.LANES), "donate_min": list(T.DONATE_MIN),
                               "build_mode": list(T.BUILD)}},
        "parity": parity,
        "host_s": {"prepare": round(t_prep, 2)},
    }
    print(json.dumps(out))


def run_serve(args):
    """Poisson stream over the 207 central-node patterns through the server.

    plain  : the same stream, one query at a time (one launch sequence each)
             on the resident graph, every query run until done or its deadline
    served : the serving loop -- lanes running concurrently, footprint
             admission, and a pilot-based prediction that refuses plans
             predicted above the deadline before they occupy a lane
    cold   : a sample of the stream the way the reference serves a query
             (prepare the graph for the query, upload, count, free)
    """
    import threading

    import numpy as np

    import beep_amd as B
    from beep_amd import patterns as P
    from beep_amd import workloads as W

    name = args.workload or "chunglu3m"
    wl = W.WORKLOADS[name]
    t0 = time.time()
    g = W.make_graph(name)
    print(f"[bench] generated {name} in {time.time() - t0:.1f} s ({g.m // 2} edges)", file=sys.stderr, flush=True)
    pats = [p for p in P.central207() if p[0] <= args.max_k]
    rng = np.random.default_rng(args.seed)
    N = args.queries
    picks = rng.integers(0, len(pats), size=N)
    sgm = rng.random(N) < 0.2
    plans = {}
    for i in range(N):
        key = (int(picks[i]), bool(sgm[i]))
        if key not in plans:
            s, d = P.directed(pats[key[0]][1])
            plans[key] = B.Plan.from_edges(s, d, B.SGM if key[1] else B.SGC)
    stream = [(int(picks[i]), bool(sgm[i])) for i in range(N)]
    cap = 1_000_000
    dl = args.deadline_ms

    def submit(srv, key):
        if key[1]:
            return srv.submit(plans[key], emit=B.EMIT_BUFFER, match_cap=cap, deadline_ms=dl)
        return srv.submit(plans[key], deadline_ms=dl)

    def drain(srv, n, tag):
        out = {}
        while len(out) < n:
            r = srv.poll(timeout_ms=1000)
            if r is not None:
                r.pop("matches", None)
                out[r["ticket"]] = r
                if len(out) % 5 == 0:
                    print(f"[bench] {tag} {len(out)}/{n}", file=sys.stderr, flush=True)
        return out

    # plain: one query at a time on the resident graph
    plain_srv = B.Server(g, lanes=1)
    t0 = time.time()
    for p in plans.values():
        plain_srv.prepare(p)
    t_warm = time.time() - t0
    print(f"[bench] warmed {len(plans)} plans in {t_warm:.1f} s", file=sys.stderr, flush=True)
    plain = []
    t0 = time.time()
    for i, key in enumerate(stream):
        submit(plain_srv, key)
        r = plain_srv.poll()
        r.pop("matches", None)
        plain.append(r)
        print(f"[bench] plain {i + 1}/{N}: status {r['status']} {r['latency_ms']:.1f} ms", file=sys.stderr,
              flush=True)
    plain_wall = time.time() - t0
    plain_srv.close()
    svc = [r["latency_ms"] for r in plain]
    plain_ok = [i for i in range(N) if plain[i]["status"] == 0]
    plain_cap = N / (sum(svc) / 1000.0)
    plain_done_rate = len(plain_ok) / (sum(svc) / 1000.0)
    qps = args.qps or 0.8 * plain_cap
    gaps = rng.exponential(1.0 / qps, size=N)
    arrivals = np.cumsum(gaps) - gaps[0]
    # plain path under the same arrivals: a FIFO single server with the measured service times
    fin, plain_lat = 0.0, []
    for a, sv in zip(arrivals * 1000.0, svc):
        fin = max(fin, a) + sv
        plain_lat.append(fin - a)

    srv = B.Server(g, lanes=args.serve_lanes, budget_ms=dl)
    for p in plans.values():
        srv.prepare(p)
    order = {}

    def producer():
        base = time.time()
        for i, key in enumerate(stream):
            wait = base + arrivals[i] - time.time()
            if wait > 0:
                time.sleep(wait)
            order[submit(srv, key)] = i

    t0 = time.time()
    th = threading.Thread(target=producer)
    th.start()
    results = drain(srv, N, "served")
    th.join()
    served_wall = time.time() - t0
    st = srv.stats()
    srv.close()
    # capacity: a fresh server (empty prediction cache), the whole stream submitted at once
    srv = B.Server(g, lanes=args.serve_lanes, budget_ms=dl)
    for p in plans.values():
        srv.prepare(p)
    t0 = time.time()
    for key in stream:
        submit(srv, key)
    burst = drain(srv, N, "burst")
    burst_wall = time.time() - t0
    srv.close()
    # cold: the reference's way, on a few queries of the stream
    cold = []
    for key in stream[:args.cold_sample]:
        t0 = time.time()
        pp = B.Prepared(g, plans[key])
        pp.upload(0)
        pp.count(emit=B.EMIT_NONE, deadline_ms=dl)
        del pp
        cold.append(time.time() - t0)
    print(f"[bench] cold sample {cold}", file=sys.stderr, flush=True)

    by_idx = {order[t]: r for t, r in results.items()}
    same = all((by_idx[i]["count"], by_idx[i]["hash_sum"]) == (plain[i]["count"], plain[i]["hash_sum"])
               for i in range(N) if plain[i]["status"] == 0 and by_idx[i]["status"] == 0)
    done = [by_idx[i] for i in range(N) if by_idx[i]["status"] == 0]
    lat_done = [r["latency_ms"] for r in done]
    lat_all = [r["latency_ms"] for r in by_idx.values()]
    n_status = lambda rs, c: sum(1 for r in rs if r["status"] == c)
    burst_done = n_status(burst.values(), 0)
    out = {
        "metric": "served queries/s",
        "value": len(done) / served_wall,
        "unit": "completed queries/s",
        "n_gpus": 1, "steps": N, "warmup": 0,
        "ms_per_step": served_wall * 1000.0 / N,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"{name}: {wl['kind']} n={g.n}, seed {wl['seed']}, {g.m // 2} undirected edges, "
                               f"max degree {g.max_degree}; Poisson arrivals at {qps:.3f} q/s over the "
                               f"{len(pats)} central-node patterns with 3-{args.max_k} vertices, 20% sgm with a "
                               f"{cap} tuple buffer, per-query deadline {dl} ms",
                   "queries": N, "qps": qps, "deadline_ms": dl, "lanes": args.serve_lanes},
        "served": {"p50_ms": pct(lat_done, 50), "p99_ms": pct(lat_done, 99),
                   "p50_ms_all": pct(lat_all, 50), "p99_ms_all": pct(lat_all, 99),
                   "completed": len(done), "stopped": n_status(by_idx.values(), B.ETIMEDOUT),
                   "refused": n_status(by_idx.values(), B.ERANGE),
                   "failed": sum(1 for r in by_idx.values() if r["status"] not in (0, B.ETIMEDOUT, B.ERANGE)),
                   "wall_s": served_wall, "peak_running": st["peak_running"], "pilots": st["pilots"],
                   "pool_peak_bytes": st["pool_peak"]},
        "plain": {"p50_ms": pct([plain_lat[i] for i in plain_ok], 50),
                  "p99_ms": pct([plain_lat[i] for i in plain_ok], 99),
                  "p50_ms_all": pct(plain_lat, 50), "p99_ms_all": pct(plain_lat, 99),
                  "completed": len(plain_ok), "stopped": n_status(plain, B.ETIMEDOUT),
                  "queries_per_s": plain_cap, "completed_per_s": plain_done_rate, "wall_s": plain_wall},
        "capacity": {"served_completed_per_s": burst_done / burst_wall, "served_completed": burst_done,
                     "served_wall_s": burst_wall, "plain_completed_per_s": plain_done_rate,
                     "speedup": (burst_done / burst_wall) / plain_done_rate if plain_done_rate else None},
        "predictor": {"refused_but_plain_completed": sum(1 for i in range(N) if by_idx[i]["status"] == B.ERANGE
                                                         and plain[i]["status"] == 0),
                      "admitted_but_stopped": n_status(by_idx.values(), B.ETIMEDOUT),
                      "misses": sorted({(stream[i][0], "sgm" if stream[i][1] else "sgc",
                                         round(by_idx[i]["predicted_ms"], 1), round(plain[i]["t_enum_ms"], 1))
                                        for i in range(N) if by_idx[i]["status"] == B.ERANGE
                                        and plain[i]["status"] == 0})},
        "cold": {"sample": len(cold), "s_per_query": float(np.mean(cold)) if cold else None,
                 "what": "prepare (peel + orient + roots) + upload + count + free, per query"},
        "same_result_alone_and_in_mix": same,
        "host_s": {"warm_caches": round(t_warm, 2)},
    }
    print(json.dumps(out))


def run_large(args):
    """The largest single-GPU graph: each query alone, then the ones that ran as one batch."""
    import beep_amd as B
    from beep_amd import patterns as P
    from beep_amd import workloads as W

    os.environ.setdefault("BEEP_VERBOSE", "1")
    name = args.workload or "rmat26"
    wl = W.WORKLOADS[name]
    t0 = time.time()
    g = W.make_graph(name)
    t_gen = time.time() - t0
    print(f"[bench] generated {name}: n={g.n} entries={g.m} max degree {g.max_degree} in {t_gen:.1f} s",
          file=sys.stderr, flush=True)
    dl = args.deadline_ms if args.deadline_ms > 0 else 0.0
    rows, ran = [], []
    pick = [x for x in args.large_queries.split(",") if x]
    for q, task in wl["queries"]:
        if pick and q not in pick:
            continue
        s, d = P.directed(P.read_mtx_edges(q))
        rec = {"query": q, "mode": task}
        try:
            t0 = time.time()
            pp = B.Prepared(g, B.Plan.from_edges(s, d, B.mode_of(task)))
            pp.upload(0)
            rec["prepare_s"] = round(time.time() - t0, 2)
            rec["preconditions"] = pp.preconditions()
            dsc = pp.describe()
            r = pp.count(deadline_ms=dl, profile=True)
            dom_ms = r["t_lds_class_ms"][DOMINANT_LG]
            rec.update(count=r["count"], t_enum_ms=r["t_enum_ms"], hub_ms=r["t_hub_ms"], lds_ms=r["t_lds_ms"],
                       donations=r["donations"], stopped=r["stopped"], b_alg=pp.alg_bytes,
                       balance_lds=r["balance_lds"], balance_hub=r["balance_hub"],
                       balance_cu=r["balance_cu"], cus_busy=r["cus_busy"],
                       donated_frac=r["donated_frac"],
                       dominant_kernel_ms=dom_ms, dominant_kernel_bytes=dsc.lds_class_bytes[DOMINANT_LG],
                       matches_per_s=r["count"] / (r["t_enum_ms"] / 1000.0))
            rec["dominant_kernel_gbs"] = (dsc.lds_class_bytes[DOMINANT_LG] / (dom_ms / 1000.0) / 1e9
                                          if dom_ms > 0 else None)
            if not r["stopped"]:
                ran.append(pp)
        except B.BeepError as e:
            rec["error"] = str(e)
        rows.append(rec)
        print(f"[bench] {json.dumps(rec)}", file=sys.stderr, flush=True)
    batched = None
    plain_ms = sum(r["t_enum_ms"] for r in rows if "t_enum_ms" in r and not r["stopped"])
    if len(ran) > 1 and plain_ms > args.large_batch_cap_ms:
        batched = {"skipped": f"the completed queries take {plain_ms / 1000:.1f} s alone, above "
                              f"--large-batch-cap-ms {args.large_batch_cap_ms:.0f}"}
    elif len(ran) > 1:
        try:
            b = B.Batch(ran)
            res, t_first = b.run()
            res, t = b.run()
            batched = {"queries": [p.plan.k for p in ran], "ms": t, "first_launch_ms": t_first,
                       "plain_ms": plain_ms,
                       "counts": [r["count"] for r in res]}
        except B.BeepError as e:
            batched = {"error": str(e)}
        print(f"[bench] batched {batched}", file=sys.stderr, flush=True)
    done = [r for r in rows if "t_enum_ms" in r and not r["stopped"]]
    tot_ms = sum(r["t_enum_ms"] for r in done)
    out = {
        "metric": "subgraph matches/s",
        "value": sum(r["count"] for r in done) / (tot_ms / 1000.0) if tot_ms else None,
        "unit": "matches/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": tot_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"{name}: R-MAT scale {wl['scale']}, {wl['samples']} samples (edge factor "
                               f"counts samples), seed {wl['seed']}, {g.m // 2} undirected edges after "
                               f"dedup, max degree {g.max_degree}",
                   "queries": rows, "deadline_ms": dl},
        "roofline": None,
        "batched": batched,
        "host_s": {"generate": round(t_gen, 1)},
    }
    dom = [r for r in rows if r.get("dominant_kernel_ms")]
    dms = sum(r["dominant_kernel_ms"] for r in dom)
    if dms > 0:
        ach = sum(r["dominant_kernel_bytes"] for r in dom) / (dms / 1000.0) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": DOMINANT_KERNEL, "achieved": ach, "peak": PEAK_HBM_GBS,
                           "unit": "GB/s", "frac": ach / PEAK_HBM_GBS, "traffic": None}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
