"""GPU parity: libbeep.so (HIP kernels through the C ABI) against the oracle.

Bit-exact counts and tuple-multiset hashes (and tuple sets for small cases) on
identical inputs.  Every case prints its name before touching the GPU.
"""
import numpy as np
import pytest

import beep_amd as B
import oracle_py as O
from beep_amd import patterns as P

pytestmark = pytest.mark.gpu

NAMED_FAST = ["triangle", "wedge", "tailed_triangle", "diamond", "4clique", "4star", "w5", "5clique"]


def say(*a):
    print("[case]", *a, flush=True)


def oracle_graph(g):
    rp, col = g.csr()
    src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    return O.Graph(src, col)


def run_pair(g, og, name_or_edges, mode, emit=B.EMIT_HASH, lanes=0, nthreads=0):
    edges = P.read_mtx_edges(name_or_edges) if isinstance(name_or_edges, str) else name_or_edges
    s, d = P.directed(edges)
    bp = B.Plan.from_edges(s, d, mode)
    op = O.plan(s, d, mode)
    pp = B.Prepared(g, bp)
    pp.upload(0)
    r = pp.count(emit=emit, lanes=lanes)
    o = O.Prepared(og, op).count(nthreads=nthreads, want_hash=emit != B.EMIT_NONE)
    return r, o, pp


def er_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    a = np.triu(rng.random((n, n)) < p, 1)
    u, v = np.nonzero(a)
    pairs = sorted(set(zip(u.tolist(), v.tolist())) | set(zip(v.tolist(), u.tolist())))
    return B.Graph.from_edges([x for x, _ in pairs], [y for _, y in pairs])


@pytest.mark.parametrize("mode", [B.SGM, B.SGC])
@pytest.mark.parametrize("name", NAMED_FAST)
def test_er_small_named(name, mode):
    g = er_graph(70, 0.25, 11)
    og = oracle_graph(g)
    say("er70", name, mode)
    r, o, _ = run_pair(g, og, name, mode)
    assert r["count"] == o["count"]
    assert (r["hash_sum"], r["hash_xor"]) == (o["hash_sum"], o["hash_xor"])


@pytest.fixture(scope="module")
def cl_graph():
    g = B.Graph.chung_lu(20000, 60000, 2.2, 5.0, 7)
    return g, oracle_graph(g)


@pytest.mark.parametrize("mode", [B.SGM, B.SGC])
@pytest.mark.parametrize("name", NAMED_FAST)
def test_chung_lu_hubs_named(cl_graph, name, mode):
    g, og = cl_graph
    say("chunglu20k", name, mode)
    r, o, pp = run_pair(g, og, name, mode)
    if name in ("tailed_triangle", "diamond", "4star", "w5"):
        assert pp.arrays()["n_hub_roots"] > 0  # bitmaps in HBM are exercised
    assert r["count"] == o["count"], (r, o)
    assert (r["hash_sum"], r["hash_xor"]) == (o["hash_sum"], o["hash_xor"])


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
def test_lane_widths_agree(cl_graph, lanes):
    g, og = cl_graph
    say("lanes", lanes)
    r, o, _ = run_pair(g, og, "diamond", B.SGM, lanes=lanes)
    assert r["count"] == o["count"]
    assert (r["hash_sum"], r["hash_xor"]) == (o["hash_sum"], o["hash_xor"])


@pytest.mark.parametrize("name,mode", [("diamond", B.SGM), ("w5", B.SGM), ("tailed_triangle", B.SGM),
                                       ("4star", B.SGM), ("w5", B.SGC)])
@pytest.mark.parametrize("donate_min", [0, 2, 16])
def test_hub_donation(cl_graph, name, mode, donate_min):
    """Donated level-2 subtrees (tiny thresholds force many) give the same
    counts and tuple hashes as walking every hub item on one wave."""
    g, og = cl_graph
    say("donation", name, mode, donate_min)
    s, d = P.directed(P.read_mtx_edges(name))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, mode))
    pp.upload(0)
    o = O.Prepared(og, O.plan(s, d, mode)).count(want_hash=True)
    if donate_min == 0:
        r = pp.count(emit=B.EMIT_HASH, donate=False)
        assert r["donations"] == 0
    else:
        r = pp.count(emit=B.EMIT_HASH, donate=True, donate_min=donate_min)
        assert r["donations"] > 0
    assert r["count"] == o["count"]
    assert (r["hash_sum"], r["hash_xor"]) == (o["hash_sum"], o["hash_xor"])


@pytest.mark.parametrize("name", ["triangle", "diamond", "4clique", "tailed_triangle", "w5"])
def test_match_buffer_tuples(name):
    g = er_graph(40, 0.3, 5)
    og = oracle_graph(g)
    say("buffer", name)
    edges = P.read_mtx_edges(name)
    s, d = P.directed(edges)
    bp = B.Plan.from_edges(s, d, B.SGM)
    pp = B.Prepared(g, bp)
    pp.upload(0)
    o = O.Prepared(og, O.plan(s, d, 0)).count(want_hash=True, match_cap=1 << 20)
    r = pp.count(emit=B.EMIT_BUFFER, match_cap=1 << 20)
    assert r["count"] == o["count"] == r["written"] == o["written"]
    a = r["matches"][np.lexsort(r["matches"].T[::-1])]
    b = o["matches"][np.lexsort(o["matches"].T[::-1])]
    assert np.array_equal(a, b)


def test_match_buffer_cap():
    g = er_graph(40, 0.3, 5)
    say("buffer-cap")
    s, d = P.directed(P.read_mtx_edges("diamond"))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, B.SGM))
    pp.upload(0)
    full = pp.count(emit=B.EMIT_BUFFER, match_cap=1 << 20)
    cap = full["count"] // 3
    r = pp.count(emit=B.EMIT_BUFFER, match_cap=cap)
    assert r["count"] == full["count"] and r["written"] == cap
    allrows = {tuple(x) for x in full["matches"].tolist()}
    assert {tuple(x) for x in r["matches"].tolist()} <= allrows


def star_plus_graph(degrees, seed=3):
    """Disjoint stars of the given degrees whose leaves carry random chords, so
    that roots land exactly on the bitmap-width and LDS/HBM boundaries."""
    rng = np.random.default_rng(seed)
    edges = set()
    base = 0
    for dg in degrees:
        c = base
        leaves = list(range(base + 1, base + 1 + dg))
        for v in leaves:
            edges.add((c, v))
        for _ in range(dg * 3):
            a, b = rng.choice(leaves, 2, replace=False)
            edges.add((int(min(a, b)), int(max(a, b))))
        base += dg + 1
    pairs = sorted(edges | {(b, a) for a, b in edges})
    return B.Graph.from_edges([x for x, _ in pairs], [y for _, y in pairs])


@pytest.mark.parametrize("name,mode", [("triangle", B.SGC), ("diamond", B.SGM), ("4clique", B.SGM),
                                       ("tailed_triangle", B.SGC), ("w5", B.SGM)])
def test_degree_boundaries(name, mode):
    g = star_plus_graph([63, 64, 65, 127, 128, 129, 255, 256, 511, 512, 513, 700, 1100])
    og = oracle_graph(g)
    say("boundaries", name, mode)
    r, o, pp = run_pair(g, og, name, mode)
    assert r["count"] == o["count"]
    assert (r["hash_sum"], r["hash_xor"]) == (o["hash_sum"], o["hash_xor"])


def test_empty_root_set():
    g = er_graph(30, 0.05, 2)
    say("empty-roots")
    s, d = P.directed(P.read_mtx_edges("6clique"))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, B.SGM))
    pp.upload(0)
    assert pp.count()["count"] == 0


def test_c1_triangle_sgc():
    """Config 1: R-MAT scale 16, edge factor 16, (.57,.19,.19,.05), seed 1, triangle sgc."""
    g = B.Graph.rmat(16, 16 << 16, 0.57, 0.19, 0.19, 1)
    og = oracle_graph(g)
    say("C1 triangle sgc")
    r, o, _ = run_pair(g, og, "triangle", B.SGC, emit=B.EMIT_NONE)
    assert r["count"] == o["count"]


@pytest.mark.parametrize("name", ["4clique", "diamond"])
def test_c1_sgm_hash(name):
    g = B.Graph.rmat(16, 16 << 16, 0.57, 0.19, 0.19, 1)
    og = oracle_graph(g)
    say("C1", name, "sgm hash")
    r, o, _ = run_pair(g, og, name, B.SGM)
    assert r["count"] == o["count"]
    assert (r["hash_sum"], r["hash_xor"]) == (o["hash_sum"], o["hash_xor"])


def test_golden_brute_cases():
    """Every tiny-graph case of tests/golden/brute.json (216 patterns x 8 graphs,
    sgm and sgc) through the HIP path equals the brute-force expectation."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "brute.json")))
    graphs = {}
    for name, v in gold["graphs"].items():
        e = sorted({(a, b) for a, b in v["edges"]} | {(b, a) for a, b in v["edges"]})
        graphs[name] = B.Graph.from_edges([a for a, _ in e], [b for _, b in e])
    plans = {}
    bad = []
    say("golden brute cases", len(gold["cases"]))
    for c in gold["cases"]:
        s, d = P.directed(gold["patterns"][c["pattern"]]["edges"])
        for mode, key in ((B.SGM, "sgm"), (B.SGC, "sgc")):
            pk = (c["pattern"], mode)
            if pk not in plans:
                plans[pk] = B.Plan.from_edges(s, d, mode)
            pp = B.Prepared(graphs[c["graph"]], plans[pk])
            pp.upload(0)
            got = pp.count()["count"]
            if got != c[key]:
                bad.append((c["graph"], c["pattern"], key, got, c[key]))
    assert not bad, bad[:10]


@pytest.mark.parametrize("mode", [B.SGM, B.SGC])
def test_batch_graph_matches_single_queries(cl_graph, mode):
    """One captured graph over 7 queries gives each query's single-launch result,
    on every replay, with eager launches interleaved between replays."""
    g, og = cl_graph
    names = ["triangle", "tailed_triangle", "diamond", "4clique", "4star", "w5", "5clique"]
    say("batch", mode)
    preps = []
    for name in names:
        s, d = P.directed(P.read_mtx_edges(name))
        pp = B.Prepared(g, B.Plan.from_edges(s, d, mode))
        pp.upload(0)
        preps.append(pp)
    single = [pp.count(emit=B.EMIT_HASH) for pp in preps]
    batch = B.Batch(preps, emit=B.EMIT_HASH)
    for rep in range(3):
        res, t = batch.run()
        for name, r, s1 in zip(names, res, single):
            assert (r["count"], r["hash_sum"], r["hash_xor"]) == (s1["count"], s1["hash_sum"], s1["hash_xor"]), (rep, name)
        preps[rep].count()  # eager work between replays
    s, d = P.directed(P.read_mtx_edges("diamond"))
    o = O.Prepared(og, O.plan(s, d, mode)).count(want_hash=True)
    assert single[2]["count"] == o["count"]


@pytest.mark.parametrize("name,mode", [("w5", B.SGM), ("diamond", B.SGM), ("4clique", B.SGC)])
def test_profiled_run_counts_the_same(cl_graph, name, mode):
    g, og = cl_graph
    say("profile", name, mode)
    s, d = P.directed(P.read_mtx_edges(name))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, mode))
    pp.upload(0)
    a = pp.count(emit=B.EMIT_HASH)
    b = pp.count(emit=B.EMIT_HASH, profile=True)
    assert (a["count"], a["hash_sum"], a["hash_xor"]) == (b["count"], b["hash_sum"], b["hash_xor"])
    assert 0.0 < b["balance_lds"] <= 1.0
    assert 0.0 <= b["balance_hub"] <= 1.0 and 0.0 <= b["donated_frac"] <= 1.0
    if pp.arrays()["n_hub_roots"]:
        assert b["balance_hub"] > 0.0


@pytest.mark.parametrize("build", [B.BUILD_SCAN, B.BUILD_SEARCH])
@pytest.mark.parametrize("name,mode", [("triangle", B.SGC), ("diamond", B.SGM), ("4clique", B.SGM),
                                       ("tailed_triangle", B.SGM), ("w5", B.SGM), ("5clique", B.SGC)])
def test_bitmap_build_modes(cl_graph, name, mode, build):
    """Streaming the larger-key ranges and binary-searching the root's
    neighbours in them find the same adjacent pairs."""
    g, og = cl_graph
    say("build mode", build, name, mode)
    s, d = P.directed(P.read_mtx_edges(name))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, mode))
    pp.upload(0)
    r = pp.count(emit=B.EMIT_HASH, build_mode=build)
    o = O.Prepared(og, O.plan(s, d, mode)).count(want_hash=True)
    assert (r["count"], r["hash_sum"], r["hash_xor"]) == (o["count"], o["hash_sum"], o["hash_xor"])


@pytest.mark.parametrize("build", [B.BUILD_SCAN, B.BUILD_SEARCH])
def test_bitmap_build_modes_small_graphs(build):
    g = er_graph(70, 0.25, 11)
    og = oracle_graph(g)
    for name in ("triangle", "diamond", "4clique", "w5"):
        for mode in (B.SGM, B.SGC):
            say("build mode er70", build, name, mode)
            s, d = P.directed(P.read_mtx_edges(name))
            pp = B.Prepared(g, B.Plan.from_edges(s, d, mode))
            pp.upload(0)
            r = pp.count(emit=B.EMIT_HASH, build_mode=build)
            o = O.Prepared(og, O.plan(s, d, mode)).count(want_hash=True)
            assert (r["count"], r["hash_sum"], r["hash_xor"]) == (o["count"], o["hash_sum"], o["hash_xor"])


def _clique(k):
    return [(a, b) for a in range(k) for b in range(a + 1, k)]


def _wheel(k):
    return [(0, i) for i in range(1, k)] + [(i, i % (k - 1) + 1) for i in range(1, k)]


_CEN10 = [(0, i) for i in range(1, 10)] + [(1, 2), (1, 5), (2, 6), (3, 4), (3, 8), (4, 9), (5, 7), (6, 9),
                                           (7, 8), (2, 9), (1, 8)]


@pytest.mark.parametrize("mode", [B.SGM, B.SGC])
@pytest.mark.parametrize("name,edges,n,p", [("8clique", _clique(8), 40, 0.8), ("9clique", _clique(9), 40, 0.8),
                                            ("10clique", _clique(10), 40, 0.8), ("w9", _wheel(9), 30, 0.5),
                                            ("central10", _CEN10, 30, 0.5)])
def test_largest_queries(name, edges, n, p, mode):
    """Queries of 8-10 vertices (the reference's DEPTH = 10) against the oracle."""
    g = er_graph(n, p, 7)
    og = oracle_graph(g)
    say("largest queries", name, mode)
    r, o, _ = run_pair(g, og, edges, mode)
    assert r["count"] == o["count"]
    assert (r["hash_sum"], r["hash_xor"]) == (o["hash_sum"], o["hash_xor"])


def test_deep_pattern_on_hubs(cl_graph):
    """A 7-vertex wheel (4 candidate levels per wave) on the hub path."""
    g, og = cl_graph
    say("deep pattern hubs w7")
    r, o, pp = run_pair(g, og, _wheel(7), B.SGC)
    assert pp.arrays()["n_hub_roots"] > 0
    assert r["count"] == o["count"]
    assert (r["hash_sum"], r["hash_xor"]) == (o["hash_sum"], o["hash_xor"])


# Central-node patterns (index into patterns.central207()) whose last level
# does not intersect with the row of the second-to-last position: count-only
# runs sum that level in closed form.  4 and 5 vertices, with and without the
# symmetry bound between the last two positions, and one of 6 vertices.
TAIL_SUMMED = [2, 3, 4, 7, 8, 9, 10, 12, 13, 15, 20]


def _central(i):
    return P.central207()[i][1]


def test_tail_summed_plans_are_the_ones_listed():
    for i in TAIL_SUMMED:
        p = B.Plan.from_edges(*P.directed(_central(i)), B.SGM)
        assert p.width >= 4 and (p.width - 2) not in p.qedge[p.width - 1][1:], i


@pytest.mark.parametrize("idx", TAIL_SUMMED)
def test_summed_last_level_on_hubs(cl_graph, idx):
    """Count-only runs (summed last level, LDS batches and hub items) equal the oracle."""
    g, og = cl_graph
    say("summed last level", idx)
    r, o, pp = run_pair(g, og, _central(idx), B.SGM, emit=B.EMIT_NONE)
    assert pp.arrays()["n_hub_roots"] > 0
    assert r["count"] == o["count"], (r["count"], o["count"])


@pytest.mark.parametrize("idx", [7, 8, 12, 20])
@pytest.mark.parametrize("donate_min", [2, 16])
def test_summed_last_level_donated(cl_graph, idx, donate_min):
    """Deeper plans keep donating level-2 subtrees; the summed last level then
    runs on the donated ranges too."""
    g, og = cl_graph
    say("summed last level donated", idx, donate_min)
    s, d = P.directed(_central(idx))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, B.SGM))
    pp.upload(0)
    o = O.Prepared(og, O.plan(s, d, 0)).count(want_hash=False)
    r = pp.count(emit=B.EMIT_NONE, donate=True, donate_min=donate_min)
    assert r["donations"] > 0
    assert r["count"] == o["count"]


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
def test_summed_last_level_lane_widths(cl_graph, lanes):
    g, og = cl_graph
    say("summed last level lanes", lanes)
    for idx in (4, 9):
        r, o, _ = run_pair(g, og, _central(idx), B.SGM, emit=B.EMIT_NONE, lanes=lanes)
        assert r["count"] == o["count"], idx


@pytest.mark.parametrize("idx", [3, 4, 9, 13])
def test_summed_last_level_degree_boundaries(idx):
    g = star_plus_graph([63, 64, 65, 127, 128, 129, 255, 256, 511, 512, 513, 700, 1100])
    og = oracle_graph(g)
    say("summed last level boundaries", idx)
    r, o, _ = run_pair(g, og, _central(idx), B.SGM, emit=B.EMIT_NONE)
    assert r["count"] == o["count"]


# Central-node patterns whose last three levels pair up (KPlan::pair): W5
# (hub + 4-cycle, index 14) and four of 6-7 vertices; tests/test_pair_sum.py
# lists every qualifying plan and checks the identity the kernels use.
PAIR_SUMMED = [14, 48, 165, 166, 204]


@pytest.mark.parametrize("idx", PAIR_SUMMED)
@pytest.mark.parametrize("mode", [B.SGM, B.SGC])
def test_paired_levels_on_hubs(cl_graph, idx, mode):
    """Count-only runs (paired last levels on LDS batches and hub items)
    equal the oracle."""
    g, og = cl_graph
    say("paired levels", idx, mode)
    r, o, pp = run_pair(g, og, _central(idx), mode, emit=B.EMIT_NONE)
    assert pp.arrays()["n_hub_roots"] > 0
    assert r["count"] == o["count"], (r["count"], o["count"])


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
def test_paired_levels_lane_widths(cl_graph, lanes):
    g, og = cl_graph
    say("paired levels lanes", lanes)
    for idx in (14, 48):
        r, o, _ = run_pair(g, og, _central(idx), B.SGM, emit=B.EMIT_NONE, lanes=lanes)
        assert r["count"] == o["count"], idx


@pytest.mark.parametrize("idx", [14, 204])
def test_paired_levels_degree_boundaries(idx):
    g = star_plus_graph([63, 64, 65, 127, 128, 129, 255, 256, 511, 512, 513, 700, 1100])
    og = oracle_graph(g)
    say("paired levels boundaries", idx)
    r, o, _ = run_pair(g, og, _central(idx), B.SGM, emit=B.EMIT_NONE)
    assert r["count"] == o["count"]


@pytest.mark.parametrize("donate_min", [2, 16])
def test_paired_levels_deeper_plans_donate(cl_graph, donate_min):
    """A 6-vertex paired plan still donates level-2 subtrees; the paired sum
    then runs on the donated ranges too."""
    g, og = cl_graph
    say("paired levels donated", donate_min)
    s, d = P.directed(_central(48))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, B.SGM))
    pp.upload(0)
    o = O.Prepared(og, O.plan(s, d, 0)).count(want_hash=False)
    r = pp.count(emit=B.EMIT_NONE, donate=True, donate_min=donate_min)
    assert r["donations"] > 0
    assert r["count"] == o["count"]


def test_golden_quirk_cases():
    """Root selection where it decides the count (tests/golden/quirks.json:
    a vertex at the root window's upper bound, the ascending head skip with
    b >= 1, regular graphs with b = 0): the HIP path equals the brute force."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "quirks.json")))
    graphs = {}
    for name, v in gold["graphs"].items():
        e = sorted({(a, b) for a, b in v["edges"]} | {(b, a) for a, b in v["edges"]})
        graphs[name] = B.Graph.from_edges([a for a, _ in e], [b for _, b in e])
    bad = []
    say("golden quirk cases", len(gold["cases"]))
    for c in gold["cases"]:
        s, d = P.directed(gold["patterns"][c["pattern"]])
        pp = B.Prepared(graphs[c["graph"]], B.Plan.from_edges(s, d, c["mode"]))
        pp.upload(0)
        for emit in (B.EMIT_NONE, B.EMIT_HASH):
            got = pp.count(emit=emit)["count"]
            if got != c["count"]:
                bad.append((c["graph"], c["pattern"], c["mode"], emit, got, c["count"]))
    assert not bad, bad[:10]
