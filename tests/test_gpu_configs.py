"""GPU counts on the full-size workload graphs against the oracle's counts
(tests/golden/<workload>_counts.json, written by make_config_counts.py), with
the root-window preconditions P1 / P2 printed for every case."""
import json
import os

import pytest

import beep_amd as B
from beep_amd import patterns as P
from beep_amd import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


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
            if q in g:
                out.append((name, q, task, g[q]))
    return out


_graphs = {}


def graph(name):
    if name not in _graphs:
        _graphs.clear()
        _graphs[name] = W.make_graph(name)
    return _graphs[name]


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name,q,task,want", cases())
def test_full_size_counts(name, q, task, want):
    print("[case]", name, q, task, flush=True)
    g = graph(name)
    s, d = P.directed(P.read_mtx_edges(q))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, B.mode_of(task)))
    pc = pp.preconditions()
    print("[preconditions]", name, q, task, json.dumps(pc), flush=True)
    pp.upload(0)
    r = pp.count()
    assert r["stopped"] == 0
    assert r["count"] == want


@pytest.mark.timeout(300)
def test_c3_batched_graph_full_size():
    """The 8 named patterns of rmat22cut (sgc) in one captured hipGraph, W5
    included, each equal to the oracle's full-size count."""
    name = "rmat22cut"
    want = golden(name)
    g = graph(name)
    queries = W.WORKLOADS[name]["queries"]
    print("[case] C3 batched", [q for q, _ in queries], flush=True)
    preps = []
    for q, task in queries:
        s, d = P.directed(P.read_mtx_edges(q))
        pp = B.Prepared(g, B.Plan.from_edges(s, d, B.mode_of(task)))
        pp.upload(0)
        preps.append(pp)
    res, t = B.Batch(preps).run()
    print("[batched]", f"{t:.1f} ms", flush=True)
    got = {q: r["count"] for (q, _), r in zip(queries, res)}
    assert all(r["stopped"] == 0 for r in res)
    assert got == {q: want[q] for q, _ in queries}
