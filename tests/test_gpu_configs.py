"""GPU counts on the full-size workload graphs against the oracle's counts
(tests/golden/<workload>_counts.json, written by make_config_counts.py)."""
import json
import os

import pytest

import beep_amd as B
from beep_amd import patterns as P
from beep_amd import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SLOW = {"w5"}  # tens of seconds on rmat22cut; covered by bench.py --mode batched


def golden(name):
    with open(os.path.join(HERE, "golden", f"{name}_counts.json")) as f:
        return json.load(f)


def cases():
    out = []
    for name in ("rmat16", "rmat22cut", "chunglu1m"):
        path = os.path.join(HERE, "golden", f"{name}_counts.json")
        if not os.path.exists(path):
            continue
        g = golden(name)
        for q, task in W.WORKLOADS[name]["queries"]:
            if q in g and q not in SLOW:
                out.append((name, q, task, g[q]))
    return out


_graphs = {}


@pytest.mark.parametrize("name,q,task,want", cases())
def test_full_size_counts(name, q, task, want):
    print("[case]", name, q, task, flush=True)
    if name not in _graphs:
        _graphs.clear()
        _graphs[name] = W.make_graph(name)
    g = _graphs[name]
    s, d = P.directed(P.read_mtx_edges(q))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, B.mode_of(task)))
    pp.upload(0)
    r = pp.count()
    assert r["stopped"] == 0
    assert r["count"] == want
