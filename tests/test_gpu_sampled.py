"""GPU counts over root subsets (beep_prep_subset) against the oracle on the
same roots, where the oracle cannot enumerate every root of the graph:

* random root subsets of a 20k-vertex Chung-Lu graph (oracle run here);
* the serving graph chunglu3m (Chung-Lu n = 3,072,441, gamma 2.0, about 117M
  undirected edges) at full size, over the stratified root samples whose
  oracle counts and tuple hashes tests/golden/make_sampled_counts.py wrote;
* queries run through the server on chunglu3m against the same queries alone;
* the scale-26 R-MAT graph (67.1M vertices, 1.76e9 undirected edges) at full
  size: the oracle's enumeration runs here on the product's prepared arrays
  (its single-threaded preparation would take too long at this size; the
  preparation is checked against the oracle's on the smaller graphs).

Each case prints its name and the root-window preconditions P1 / P2 first.
"""
import json
import os
import time

import numpy as np
import pytest

import beep_amd as B
import oracle_py as O
from beep_amd import patterns as P
from beep_amd import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print("[case]", *a, flush=True)


def plan_of(name, mode):
    s, d = P.directed(P.read_mtx_edges(name))
    return B.Plan.from_edges(s, d, mode)


def oracle_graph(g):
    rp, col = g.csr()
    src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    return O.Graph(src, col)


@pytest.fixture(scope="module")
def cl20k():
    g = B.Graph.chung_lu(20000, 60000, 2.2, 5.0, 7)
    return g, oracle_graph(g)


@pytest.mark.parametrize("name,mode", [("diamond", B.SGM), ("w5", B.SGM), ("4clique", B.SGC),
                                       ("tailed_triangle", B.SGM), ("5clique", B.SGM)])
def test_random_subsets_match_oracle(cl20k, name, mode):
    g, og = cl20k
    say("subset chunglu20k", name, mode)
    s, d = P.directed(P.read_mtx_edges(name))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, mode))
    op = O.Prepared(og, O.plan(s, d, mode))
    n = pp.n_roots
    rng = np.random.default_rng(3)
    for frac in (0.02, 0.3):
        ranks = np.sort(rng.choice(n, size=max(1, int(n * frac)), replace=False)).astype(np.uint64)
        sub = pp.subset(ranks)
        sub.upload(0)
        r = sub.count(emit=B.EMIT_HASH)
        o = op.count_ranks(ranks, want_hash=True)
        assert (r["count"], r["hash_sum"], r["hash_xor"]) == (o["count"], o["hash_sum"], o["hash_xor"])


# ---------------------------------------------------------------- chunglu3m

def c4_golden():
    path = os.path.join(HERE, "golden", "chunglu3m_sampled.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


_c4 = {}


def c4_graph():
    if "g" not in _c4:
        t = time.time()
        _c4["g"] = W.make_graph("chunglu3m")
        print(f"[c4] generated in {time.time() - t:.1f} s", flush=True)
    return _c4["g"]


@pytest.mark.timeout(300)
@pytest.mark.parametrize("key", sorted(c4_golden()))
def test_c4_sampled_roots_match_oracle(key):
    want = c4_golden()[key]
    say("C4 chunglu3m sampled", key, f"stride {want['stride']}")
    g = c4_graph()
    pp = B.Prepared(g, plan_of(want["query"], B.mode_of(want["task"])))
    print("[preconditions] chunglu3m", key, json.dumps(pp.preconditions()), flush=True)
    assert pp.n_roots == want["nroots"]
    ranks = np.arange(want.get("first", 0) + want["offset"], want["nroots"], want["stride"], dtype=np.uint64)
    assert len(ranks) == want["sampled"]
    sub = pp.subset(ranks)
    sub.upload(0)
    r = sub.count(emit=B.EMIT_HASH)
    assert r["stopped"] == 0
    assert (r["count"], r["hash_sum"], r["hash_xor"]) == (want["count"], want["hash_sum"], want["hash_xor"])


C4_SERVED = [("triangle", B.SGC), ("wedge", B.SGM), ("tailed_triangle", B.SGC), ("4star", B.SGC),
             ("diamond", B.SGC)]


@pytest.mark.timeout(300)
def test_c4_served_equals_alone():
    """Queries on the full chunglu3m graph submitted to the server back to
    back (they overlap on the device) give the counts and tuple hashes they
    give alone."""
    say("C4 chunglu3m served vs alone", C4_SERVED)
    g = c4_graph()
    ref = []
    for name, mode in C4_SERVED:
        pp = B.Prepared(g, plan_of(name, mode))
        pp.upload(0)
        ref.append(pp.count(emit=B.EMIT_HASH))
        print("[alone]", name, mode, ref[-1]["count"], f"{ref[-1]['t_enum_ms']:.1f} ms", flush=True)
        del pp
    srv = B.Server(g, lanes=4)
    plans = [plan_of(n, m) for n, m in C4_SERVED]
    for p in plans:
        srv.prepare(p)
    tickets = {srv.submit(p, emit=B.EMIT_HASH): i for i, p in enumerate(plans)}
    got = {}
    while len(got) < len(tickets):
        r = srv.poll(timeout_ms=120000)
        assert r is not None, "server made no progress"
        got[r["ticket"]] = r
    st = srv.stats()
    srv.close()
    for t, i in tickets.items():
        r = got[t]
        assert r["status"] == 0 and r["stopped"] == 0, (C4_SERVED[i], r)
        assert (r["count"], r["hash_sum"], r["hash_xor"]) == \
            (ref[i]["count"], ref[i]["hash_sum"], ref[i]["hash_xor"]), C4_SERVED[i]
    assert st["completed"] == len(plans) and st["failed"] == 0
    print("[served] peak_running", st["peak_running"], flush=True)


# ---------------------------------------------------------------- rmat26

C5_QUERIES = [("4clique", "sgc", 1 << 20), ("w5", "sgc", 1024)]  # (query, task, max sampled degree)
C5_SAMPLE = 2000  # roots per query


@pytest.mark.timeout(900)
def test_c5_sampled_roots_match_oracle():
    """rmat26 at full size: per query, a stratified sample of C5_SAMPLE roots
    of degree at most the query's cap (the heaviest roots alone would take
    the oracle hours), counted on the GPU and by the oracle."""
    say("C5 rmat26 sampled", C5_QUERIES)
    t = time.time()
    g = W.make_graph("rmat26")
    print(f"[c5] generated in {time.time() - t:.1f} s: n {g.n}, m {g.m}", flush=True)
    for q, task, dcap in C5_QUERIES:
        mode = B.mode_of(task)
        t = time.time()
        pp = B.Prepared(g, plan_of(q, mode))
        print(f"[c5] {q} {task}: prepared in {time.time() - t:.1f} s", flush=True)
        print("[preconditions] rmat26", q, task, json.dumps(pp.preconditions()), flush=True)
        a = pp.arrays(copy=False)
        rp, roots = a["rowptr"], a["roots"]
        deg = (rp[roots.astype(np.int64) + 1] - rp[roots.astype(np.int64)]).astype(np.int64)
        first = int(np.searchsorted(-deg, -dcap, side="left"))  # roots sorted by degree descending
        stride = max(1, (len(roots) - first) // C5_SAMPLE)
        ranks = np.arange(first + stride // 2, len(roots), stride, dtype=np.uint64)
        sub = pp.subset(ranks)
        sub.upload(0)
        r = sub.count(emit=B.EMIT_HASH)
        print(f"[c5] {q}: GPU {r['count']} over {len(ranks)} roots from rank {first} "
              f"(degree <= {dcap}), {r['t_enum_ms']:.1f} ms", flush=True)
        sp, dp = P.directed(P.read_mtx_edges(q))
        view = O.PreparedView(O.plan(sp, dp, mode), rp, a["col"], a["ocol"], a["nlow"], roots, a["asc"])
        t = time.time()
        o = view.count_ranks(ranks, nthreads=min(16, os.cpu_count() or 1), want_hash=True)
        print(f"[c5] {q}: oracle {o['count']} in {time.time() - t:.1f} s", flush=True)
        assert r["stopped"] == 0
        assert (r["count"], r["hash_sum"], r["hash_xor"]) == (o["count"], o["hash_sum"], o["hash_xor"])
        del view, sub, a, pp
