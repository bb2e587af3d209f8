"""Writes tests/golden/<workload>_counts.json: counts of the full-size
workload graphs computed by the oracle (the CPU restatement in oracle/).

These pin the GPU path at the configurations' real sizes.  The oracle is slow
on the heavy patterns (hours for W5 on rmat22cut with 8 threads), so queries
can be chosen and the file is merged, not overwritten.

Run: python tests/golden/make_config_counts.py rmat22cut [query ...]
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


def main():
    name = sys.argv[1]
    wanted = set(sys.argv[2:])
    path = os.path.join(HERE, f"{name}_counts.json")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    g = W.make_graph(name)
    rp, col = g.csr()
    src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    og = O.Graph(src, col)
    for q, task in W.WORKLOADS[name]["queries"]:
        if wanted and q not in wanted:
            continue
        s, d = P.directed(P.read_mtx_edges(q))
        op = O.Prepared(og, O.plan(np.array(s, dtype=np.uint32), np.array(d, dtype=np.uint32), mode_of(task)))
        t = time.time()
        r = op.count(nthreads=os.cpu_count() or 1)
        out[q] = r["count"]
        print(q, task, r["count"], f"{time.time() - t:.1f} s", flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
