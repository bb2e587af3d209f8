"""GPU tests of deadlines and of the serving loop (beep_server through the C ABI).

The server's results must be the ones the same query gives alone (counts and
tuple-multiset hashes bit-exact, tuple sets equal for sgm with a buffer),
whatever else runs beside it; admission must respect the pool; deadlines must
stop a query, mark it, and leave the next run unaffected.
"""
import numpy as np
import pytest

import beep_amd as B
from beep_amd import patterns as P
from beep_amd import workloads as W

pytestmark = pytest.mark.gpu

ENOMEM = -12


def say(*a):
    print("[case]", *a, flush=True)


def plan_of(name, mode):
    s, d = P.directed(P.read_mtx_edges(name))
    return B.Plan.from_edges(s, d, mode)


@pytest.fixture(scope="module")
def cl_graph():
    return B.Graph.chung_lu(20000, 60000, 2.2, 5.0, 7)


@pytest.fixture(scope="module")
def c2_graph():
    return W.make_graph("chunglu1m")


MIX = [("triangle", B.SGC), ("tailed_triangle", B.SGM), ("diamond", B.SGC), ("diamond", B.SGM),
       ("4clique", B.SGM), ("4star", B.SGC), ("w5", B.SGM), ("5clique", B.SGC), ("wedge", B.SGM),
       ("tailed_triangle", B.SGC), ("4clique", B.SGC), ("w5", B.SGC)]


def alone(g, name, mode, emit=B.EMIT_HASH, match_cap=0):
    pp = B.Prepared(g, plan_of(name, mode))
    pp.upload(0)
    return pp.count(emit=emit, match_cap=match_cap)


@pytest.mark.parametrize("lanes", [1, 4])
def test_server_mix_matches_alone(cl_graph, lanes):
    say("server mix lanes", lanes)
    ref = [alone(cl_graph, n, m) for n, m in MIX]
    srv = B.Server(cl_graph, lanes=lanes)
    plans = [plan_of(n, m) for n, m in MIX]
    # every query twice, submitted back to back: they overlap on the device
    tickets = {}
    for rep in range(2):
        for i, p in enumerate(plans):
            tickets[srv.submit(p, emit=B.EMIT_HASH)] = i
    got = {}
    while len(got) < len(tickets):
        r = srv.poll(timeout_ms=60000)
        assert r is not None, "server made no progress"
        got[r["ticket"]] = r
    st = srv.stats()
    srv.close()
    for t, i in tickets.items():
        r = got[t]
        assert r["status"] == 0, (MIX[i], r)
        assert r["stopped"] == 0
        assert (r["count"], r["hash_sum"], r["hash_xor"]) == \
            (ref[i]["count"], ref[i]["hash_sum"], ref[i]["hash_xor"]), MIX[i]
        assert r["latency_ms"] >= r["service_ms"] >= 0
    assert st["completed"] == len(tickets) and st["failed"] == 0
    assert st["peak_running"] <= lanes
    if lanes > 1:
        assert st["peak_running"] > 1  # queries did overlap


def test_server_sgm_tuples_match_alone():
    say("server sgm tuples")
    # 2,222,936 diamonds with 5 hub roots, every tuple fits the buffer
    g = B.Graph.chung_lu(20000, 40000, 2.1, 2.0, 3)
    names = ["diamond", "triangle", "4clique"]
    srv = B.Server(g, lanes=3)
    ref, tickets = {}, {}
    for n in names:
        a = alone(g, n, B.SGM, emit=B.EMIT_BUFFER, match_cap=1 << 22)
        assert a["written"] == a["count"]
        ref[n] = a
        tickets[srv.submit(plan_of(n, B.SGM), emit=B.EMIT_BUFFER, match_cap=1 << 22)] = n
    for _ in names:
        r = srv.poll(timeout_ms=60000)
        n = tickets[r["ticket"]]
        assert r["status"] == 0 and r["written"] == r["count"] == ref[n]["count"]
        a = ref[n]["matches"]
        b = r["matches"]
        assert sorted(map(tuple, a.tolist())) == sorted(map(tuple, b.tolist())), n
    srv.close()


def test_server_pool_admission(cl_graph):
    say("server pool admission")
    # a pool that holds one hub query's scratch at a time
    probe = B.Server(cl_graph, lanes=1)
    t = probe.submit(plan_of("w5", B.SGM), emit=B.EMIT_HASH)  # the footprint of the queries below
    need = probe.poll(timeout_ms=60000)["footprint"]
    probe.close()
    assert need > 0
    srv = B.Server(cl_graph, lanes=4, pool_bytes=max(need + need // 2, 1 << 20))
    ref = alone(cl_graph, "w5", B.SGM)
    ts = [srv.submit(plan_of("w5", B.SGM), emit=B.EMIT_HASH) for _ in range(4)]
    res = [srv.poll(timeout_ms=60000) for _ in ts]
    st = srv.stats()
    srv.close()
    assert st["peak_running"] == 1          # memory, not lanes, limited admission
    assert st["pool_peak"] <= st["pool_bytes"]
    for r in res:
        assert r["status"] == 0 and (r["count"], r["hash_sum"]) == (ref["count"], ref["hash_sum"])
    # a query that can never fit is refused, the server keeps serving
    tiny = B.Server(cl_graph, lanes=2, pool_bytes=1 << 20)
    t_big = tiny.submit(plan_of("w5", B.SGM), emit=B.EMIT_HASH)
    t_small = tiny.submit(plan_of("triangle", B.SGC))
    out = {}
    for _ in range(2):
        r = tiny.poll(timeout_ms=60000)
        out[r["ticket"]] = r
    tiny.close()
    assert out[t_big]["status"] == ENOMEM
    if out[t_small]["footprint"] <= 1 << 20:
        assert out[t_small]["status"] == 0
        assert out[t_small]["count"] == alone(cl_graph, "triangle", B.SGC)["count"]


def test_deadline_stops_and_recovers(c2_graph):
    say("deadline chunglu1m diamond sgm")
    pp = B.Prepared(c2_graph, plan_of("diamond", B.SGM))
    pp.upload(0)
    full = pp.count(emit=B.EMIT_HASH)
    assert full["stopped"] == 0
    part = pp.count(emit=B.EMIT_HASH, deadline_ms=0.05)
    assert part["stopped"] == 1
    assert part["count"] < full["count"]
    again = pp.count(emit=B.EMIT_HASH)
    assert again["stopped"] == 0
    assert (again["count"], again["hash_sum"], again["hash_xor"]) == \
        (full["count"], full["hash_sum"], full["hash_xor"])


def test_server_deadline(c2_graph):
    say("server deadline chunglu1m")
    srv = B.Server(c2_graph, lanes=2)
    t_cut = srv.submit(plan_of("diamond", B.SGM), deadline_ms=0.05)
    t_ok = srv.submit(plan_of("4clique", B.SGM), emit=B.EMIT_HASH)
    out = {}
    for _ in range(2):
        r = srv.poll(timeout_ms=120000)
        out[r["ticket"]] = r
    srv.close()
    assert out[t_cut]["status"] == B.ETIMEDOUT and out[t_cut]["stopped"] == 1
    ref = alone(c2_graph, "4clique", B.SGM)
    assert out[t_ok]["status"] == 0
    assert (out[t_ok]["count"], out[t_ok]["hash_sum"]) == (ref["count"], ref["hash_sum"])


def test_server_prediction_refuses_heavy_plans(c2_graph):
    say("server prediction chunglu1m")
    srv = B.Server(c2_graph, lanes=2, budget_ms=50.0)
    # with tuple hashes every one of the sum over roots of C(d, 3) tuples
    # (d up to 29k) is walked
    heavy = plan_of("4star", B.SGM)
    light = plan_of("triangle", B.SGC)
    srv.submit(heavy, emit=B.EMIT_HASH)
    r1 = srv.poll(timeout_ms=120000)
    assert r1["status"] == B.ERANGE and r1["predicted_ms"] > 50.0
    srv.submit(light, emit=B.EMIT_HASH)
    r2 = srv.poll(timeout_ms=120000)
    ref = alone(c2_graph, "triangle", B.SGC)
    assert r2["status"] == 0 and 0.0 <= r2["predicted_ms"] <= 4 * 50.0
    assert (r2["count"], r2["hash_sum"], r2["hash_xor"]) == (ref["count"], ref["hash_sum"], ref["hash_xor"])
    pilots = srv.stats()["pilots"]
    srv.submit(heavy, emit=B.EMIT_HASH)  # refused from the cached prediction, no new pilot
    r3 = srv.poll(timeout_ms=120000)
    st = srv.stats()
    srv.close()
    assert r3["status"] == B.ERANGE and st["pilots"] == pilots and st["refused"] == 2
