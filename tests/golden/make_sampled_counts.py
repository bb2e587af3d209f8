"""Writes tests/golden/<workload>_sampled.json: oracle counts and tuple-multiset
hashes of full-size workload graphs over a stratified sample of the root list.

Some workload graphs are too large for the oracle to enumerate every root (the
reference's per-root bitmap build alone is quadratic in the root's degree), so
both sides count the same roots: the ranks r = first + offset + i * stride of
the root list R (R in processing order, degree descending), i.e. one root from
the middle of every stride of equal width, where `first` is the first rank of
degree at most the query's cap (the heaviest roots of the heavier patterns
would take the oracle hours: the reference's bitmap build is quadratic in the
root's degree and the walk of a hub can be cubic).  The GPU test rebuilds the
same ranks from (first, stride, offset, nroots) and counts them through
beep_prep_subset.

Run: python tests/golden/make_sampled_counts.py chunglu3m query:task:stride[:degree_cap] ...
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle_py as O  # noqa: E402
from beep_amd import mode_of  # noqa: E402
from beep_amd import patterns as P  # noqa: E402
from beep_amd import workloads as W  # noqa: E402


def sample_ranks(first, nroots, stride):
    return np.arange(first + stride // 2, nroots, stride, dtype=np.uint64)


def main():
    name = sys.argv[1]
    path = os.path.join(HERE, f"{name}_sampled.json")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    t = time.time()
    g = W.make_graph(name)
    rp, col = g.csr()
    print(f"generated {name} in {time.time() - t:.1f} s", flush=True)
    src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    og = O.Graph(src, col)
    del src
    for spec in sys.argv[2:]:
        parts = spec.split(":")
        q, task, stride = parts[0], parts[1], int(parts[2])
        dcap = int(parts[3]) if len(parts) > 3 else None
        s, d = P.directed(P.read_mtx_edges(q))
        t = time.time()
        op = O.Prepared(og, O.plan(np.array(s, dtype=np.uint32), np.array(d, dtype=np.uint32), mode_of(task)))
        nroots = int(op.pp.nroots)
        first = 0
        if dcap is not None:
            a = op.arrays()
            rt = a["roots"].astype(np.int64)
            deg = (a["rowptr"][rt + 1] - a["rowptr"][rt]).astype(np.int64)
            first = int(np.searchsorted(-deg, -dcap, side="left"))  # roots are degree descending
        ranks = sample_ranks(first, nroots, stride)
        t_prep = time.time() - t
        t = time.time()
        r = op.count_ranks(ranks, nthreads=os.cpu_count() or 1, want_hash=True)
        key = f"{q}:{task}"
        out[key] = dict(query=q, task=task, nroots=nroots, first=first, degree_cap=dcap, stride=stride,
                        offset=stride // 2, sampled=len(ranks), count=r["count"], hash_sum=r["hash_sum"],
                        hash_xor=r["hash_xor"])
        print(key, out[key], f"prep {t_prep:.1f} s, count {time.time() - t:.1f} s", flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
