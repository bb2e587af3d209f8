#!/usr/bin/env python3
"""Benchmark of the enumeration path on one MI355X.

Workload (one "step"): the Chung-Lu graph chunglu1m (n = 1,134,890, gamma 2.2,
seed 2; 2,983,214 undirected edges, max degree 29,277) resident in HBM; the
4-clique query and then the diamond query, each one enumeration launch
sequence in sgm mode (counting, no tuple output -- the reference's sgm).

Reported:
  value        matches/s = (count_4clique + count_diamond) / device time
  roofline     B_alg / t_enum against 8 TB/s, B_alg = sum over roots of
               4(d(u)+2) + sum over neighbours 4(d(v)+2) (peeled degrees);
               t_enum is the HIP-event time of the query's launch sequence on
               the library's stream; "traffic" is the measured HBM bytes per
               step from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
  cpu_baseline the oracle (CPU restatement of the reference enumeration,
               OpenMP over roots) on the same prepared graphs, all host cores
               up to 16, same two queries
  parity       GPU counts and tuple-multiset hashes equal to the oracle's
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
    ap.add_argument("--workload", default="chunglu1m")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--lanes", type=int, default=0)
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
    """Bytes per step from the committed FETCH_SIZE/WRITE_SIZE summary, if any."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            if ln.startswith("bytes_per_step"):
                return float(ln.split("=")[1])
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} requested; this build measures exactly one GPU "
              "(the path is single-GPU by design; see DESIGN.md for how it would shard)",
              file=sys.stderr)
        sys.exit(2)
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

    if args.once:
        for name, task, plan, pp in preps:
            r = pp.count(emit=B.EMIT_NONE, lanes=args.lanes)
            print(json.dumps({"query": name, "count": r["count"], "t_enum_ms": r["t_enum_ms"],
                              "alg_bytes": pp.alg_bytes, "donations": r["donations"]}))
        if args.calib_copy:
            calib_copy(args.calib_copy)
        return

    # correctness run with tuple hashes (untimed)
    gpu_ref = []
    for name, task, plan, pp in preps:
        r = pp.count(emit=B.EMIT_HASH, lanes=args.lanes)
        gpu_ref.append(r)

    for _ in range(args.warmup):
        for _, _, _, pp in preps:
            pp.count(emit=B.EMIT_NONE, lanes=args.lanes)
    per_query_ms = [0.0] * len(preps)
    per_query_hub = [0.0] * len(preps)
    per_query_lds = [0.0] * len(preps)
    counts = [0] * len(preps)
    wall0 = time.time()
    for _ in range(args.steps):
        for i, (_, _, _, pp) in enumerate(preps):
            r = pp.count(emit=B.EMIT_NONE, lanes=args.lanes)
            per_query_ms[i] += r["t_enum_ms"]
            per_query_hub[i] += r["t_hub_ms"]
            per_query_lds[i] += r["t_lds_ms"]
            counts[i] = r["count"]
    wall = time.time() - wall0
    K = args.steps
    dev_ms_step = sum(per_query_ms) / K
    matches = sum(counts)
    value = matches / (dev_ms_step / 1000.0)
    alg_bytes = sum(pp.alg_bytes for _, _, _, pp in preps)
    achieved = alg_bytes / (dev_ms_step / 1000.0) / 1e9
    traffic = read_traffic(args.workload)

    # CPU baseline + parity (oracle, same prepared graphs)
    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        cores = min(os.cpu_count() or 1, 16)
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
            o = op.count(nthreads=cores, want_hash=True)
            cpu_s += time.time() - t0
            cpu_matches += o["count"]
            ok = (o["count"] == gr["count"] and o["hash_sum"] == gr["hash_sum"]
                  and o["hash_xor"] == gr["hash_xor"])
            parity = parity and ok
        cpu = {"value": cpu_matches / cpu_s, "unit": "matches/s", "cores": cores, "kind": "port",
               "sample": f"full workload ({' then '.join(q for q, _ in queries)}) on the same "
                         f"prepared graphs, {cpu_s:.2f} s of enumeration, tuple hashes on"}
        parity = bool(parity) and counts == [r["count"] for r in gpu_ref]

    out = {
        "metric": "subgraph matches/s",
        "value": value,
        "unit": "matches/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / K,
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "alg_bytes_per_step": alg_bytes},
        "cpu_baseline": cpu,
        "parity": parity,
        "per_query_ms": {q: per_query_ms[i] / K for i, (q, _) in enumerate(queries)},
        "per_query_phase_ms": {q: {"hub": per_query_hub[i] / K, "lds": per_query_lds[i] / K}
                               for i, (q, _) in enumerate(queries)},
        "host_s": {"generate": round(t_gen, 3), "prepare": round(t_prep, 3),
                   "upload": round(t_upload, 3)},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
