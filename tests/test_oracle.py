"""Pins the oracle (CPU restatement) against independent expectations.

* plans of the 207 central-node patterns + named queries against the pure-Python
  planner in tests/golden (exhaustive automorphism search) and the hand-derived
  plan table of the named patterns;
* sgm / sgc counts against brute-force constraint counting and true counts
  (#monomorphisms / |Aut|) on 8 tiny graphs x 216 patterns;
* closed forms on mid-size graphs: triangles = tr(A^3)/6, k-star sums, wedge
  sgc = 0, clique counts against networkx;
* the reader's byte semantics on hand-written .bel / .mtx / .tsv files and the
  cutoff statistic on tiny degree sequences.
"""
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle_py as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "brute.json")))


def directed(edges):
    d = sorted({(a, b) for a, b in edges} | {(b, a) for a, b in edges})
    return (np.array([a for a, _ in d], dtype=np.uint32), np.array([b for _, b in d], dtype=np.uint32))


def sym_graph(n, edges):
    s, d = directed(edges)
    return O.Graph(s, d)


@pytest.mark.parametrize("name", sorted(GOLD["patterns"]))
def test_plan_matches_independent_planner(name):
    pat = GOLD["patterns"][name]
    s, d = directed(pat["edges"])
    for mode, key in ((0, "sgm"), (1, "sgc")):
        want = GOLD["plans"][name][key]
        got = O.plan(s, d, mode).lists()
        assert got["seq"] == want["seq"]
        assert got["qedge"] == want["qedge"]
        assert got["sym"] == want["sym"]
        assert got["lunmat"] == want["lunmat"]
        assert got["asc"] == want["asc"]


# Plan table derived by hand from the reference's planner rules
# (subgraph_matching.cuh:297-351, 397-505, 612-631).
HAND_PLANS = {
    "triangle": ([0, 1, 2], [0, 0, 1, 3], [0, 0, 1], [0, 0, 1, 3], [0, 0, 1], 1, 0),
    "wedge": ([0, 1, 2], [0, 0, 1, 2], [0, 0], [0, 0, 0, 1], [1], 0, 1),
    "tailed_triangle": ([0, 1, 2, 3], [0, 0, 1, 3, 4], [0, 0, 1, 0], [0, 0, 0, 1, 1], [1], 0, 1),
    "diamond": ([0, 2, 1, 3], [0, 0, 1, 3, 5], [0, 0, 1, 0, 1], [0, 0, 1, 1, 2], [0, 2], 1, 1),
    "4clique": ([0, 1, 2, 3], [0, 0, 1, 3, 6], [0, 0, 1, 0, 1, 2], [0, 0, 1, 3, 6], [0, 0, 1, 0, 1, 2], 1, 0),
    "4star": ([0, 1, 2, 3], [0, 0, 1, 2, 3], [0, 0, 0], [0, 0, 0, 1, 3], [1, 1, 2], 0, 1),
    "w5": ([0, 1, 2, 3, 4], [0, 0, 1, 3, 5, 8], [0, 0, 1, 0, 2, 0, 1, 3], [0, 0, 0, 1, 2, 4], [1, 1, 1, 2], 0, 0),
}


@pytest.mark.parametrize("name", sorted(HAND_PLANS))
def test_plan_hand_table(name):
    seq, qp, qe, sp, se, asc, lunmat_sgc = HAND_PLANS[name]
    s, d = directed(GOLD["patterns"][name]["edges"])
    p = O.plan(s, d, 1)
    k = p.k
    assert list(p.seq[:k]) == seq
    assert list(p.qedge_ptr[:k + 1]) == qp and list(p.qedge[:qp[-1]]) == qe
    assert list(p.sym_ptr[:k + 1]) == sp and list(p.sym[:sp[-1]]) == se
    assert p.asc == asc and p.lunmat == lunmat_sgc


def _case_graphs():
    return {g: sym_graph(v["n"], v["edges"]) for g, v in GOLD["graphs"].items()}


@pytest.fixture(scope="module")
def case_graphs():
    return _case_graphs()


def test_brute_force_counts(case_graphs):
    """sgm and sgc counts equal brute-force constraint counting on every case."""
    plans = {}
    bad = []
    for c in GOLD["cases"]:
        g = case_graphs[c["graph"]]
        pat = GOLD["patterns"][c["pattern"]]
        s, d = directed(pat["edges"])
        for mode, key in ((0, "sgm"), (1, "sgc")):
            pk = (c["pattern"], mode)
            if pk not in plans:
                plans[pk] = O.plan(s, d, mode)
            got = O.Prepared(g, plans[pk]).count(nthreads=1)["count"]
            if got != c[key]:
                bad.append((c["graph"], c["pattern"], key, got, c[key]))
    assert not bad, bad[:10]


def test_sgm_equals_true_counts_here():
    """On the tiny graphs the preconditions hold, so sgm counts are true counts."""
    assert all(c["sgm"] == c["true"] for c in GOLD["cases"])


def _rand_graph(n, m, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, size=(m, 2))
    e = {(int(min(x, y)), int(max(x, y))) for x, y in a if x != y}
    return sorted(e)


def _adj(n, edges):
    import scipy.sparse as sp
    r = [a for a, b in edges] + [b for a, b in edges]
    c = [b for a, b in edges] + [a for a, b in edges]
    return sp.csr_matrix((np.ones(len(r), dtype=np.int64), (r, c)), shape=(n, n))


def _count(g, edges, mode):
    s, d = directed(edges)
    return O.Prepared(g, O.plan(s, d, mode)).count()["count"]


def test_closed_forms_midsize():
    n, edges = 1500, _rand_graph(1500, 9000, 3)
    g = sym_graph(n, edges)
    A = _adj(n, edges)
    deg = np.asarray(A.sum(axis=1)).ravel()
    tri = int((A @ A).multiply(A).sum()) // 6
    assert _count(g, [(0, 1), (0, 2), (1, 2)], 0) == tri
    assert _count(g, [(0, 1), (0, 2), (1, 2)], 1) == tri
    # k-stars: sgm sum C(d, k-1); sgc sum C(d, k-2) over roots with d >= k-1
    for k in (3, 4, 5):
        star = [(0, i) for i in range(1, k)]
        sgm = sum(math.comb(int(x), k - 1) for x in deg)
        sgc = sum(math.comb(int(x), k - 2) for x in deg if x >= k - 1) if k > 3 else 0
        assert _count(g, star, 0) == sgm
        assert _count(g, star, 1) == sgc


def test_clique_counts_networkx():
    nx = pytest.importorskip("networkx")
    n, edges = 400, _rand_graph(400, 6000, 9)
    g = sym_graph(n, edges)
    G = nx.Graph(edges)
    by_k = {}
    for c in nx.enumerate_all_cliques(G):
        by_k[len(c)] = by_k.get(len(c), 0) + 1
    for k in (4, 5):
        kc = [(a, b) for a in range(k) for b in range(a + 1, k)]
        assert _count(g, kc, 0) == by_k.get(k, 0)


def test_relabel_invariance():
    n, edges = 300, _rand_graph(300, 2500, 4)
    perm = np.random.default_rng(1).permutation(n)
    e2 = sorted({(int(min(perm[a], perm[b])), int(max(perm[a], perm[b]))) for a, b in edges})
    g1, g2 = sym_graph(n, edges), sym_graph(n, e2)
    for pat in ([(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)], [(0, 1), (0, 2), (0, 3), (1, 2)]):
        assert _count(g1, pat, 0) == _count(g2, pat, 0)


def test_match_tuples_are_matches():
    n, edges = 60, _rand_graph(60, 500, 5)
    g = sym_graph(n, edges)
    es = set(edges) | {(b, a) for a, b in edges}
    pat = [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)]
    s, d = directed(pat)
    p = O.plan(s, d, 0)
    out = O.Prepared(g, p).count(match_cap=1 << 16)
    seq = list(p.seq[:p.k])
    pos = {v: i for i, v in enumerate(seq)}
    canon = set()
    for t in out["matches"]:
        for a, b in pat:
            assert (int(t[pos[a]]), int(t[pos[b]])) in es
        canon.add(frozenset(frozenset((int(t[pos[a]]), int(t[pos[b]]))) for a, b in pat))
    assert len(canon) == out["count"]  # every subgraph once


# ---------------------------------------------------------------- reader
def _write(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(data) if isinstance(data, bytes) else p.write_text(data)
    return str(p)


def test_read_bel_layout(tmp_path):
    recs = [(1, 0), (2, 0), (0, 1), (0, 2)]  # (dst, src) pairs
    data = b"".join(struct.pack("<QQQ", dst, src, 7) for dst, src in recs) + b"\x01\x02"
    s, d = O.read_edges(_write(tmp_path, "g.bel", data))
    assert s.tolist() == [0, 0, 1, 2] and d.tolist() == [1, 2, 0, 0]


def test_read_bel_truncates_to_u32(tmp_path):
    data = struct.pack("<QQQ", (1 << 32) + 5, (1 << 33) + 3, 0)
    s, d = O.read_edges(_write(tmp_path, "t.bel", data))
    assert s.tolist() == [3] and d.tolist() == [5]


def test_read_mtx_batches(tmp_path):
    body = "%%MatrixMarket matrix coordinate pattern general\n% c\n3 3 6\n0 1\n0 2\n\n1 0\n"
    s, d = O.read_edges(_write(tmp_path, "q.mtx", body))
    # the blank line ends the first batch; the next batch continues after it
    assert list(zip(s.tolist(), d.tolist())) == [(0, 1), (0, 2), (1, 0)]
    lines = "\n".join(f"{i} {i + 1}" for i in range(100))
    s, d = O.read_edges(_write(tmp_path, "b.mtx", "3 3 3\n" + lines + "\n\n5 6\n"))
    # a batch of exactly 100 lines, then a batch that fails at once: reading stops
    assert len(s) == 100


def test_read_mtx_rejects_symmetric(tmp_path):
    with pytest.raises(IOError):
        O.read_edges(_write(tmp_path, "s.mtx", "%%MatrixMarket matrix coordinate pattern symmetric\n2 2 1\n0 1\n"))


def test_read_tsv(tmp_path):
    s, d = O.read_edges(_write(tmp_path, "g.tsv", "1 0 1\n0 1 1\n"))
    assert s.tolist() == [0, 1] and d.tolist() == [1, 0]


def test_csr_rules():
    # self loop dropped, duplicate kept, n = max src + 1, unsorted groups rejected
    g = O.Graph(np.array([0, 0, 0, 1, 1, 2], dtype=np.uint32), np.array([0, 1, 1, 0, 0, 0], dtype=np.uint32))
    rp, col = g.arrays()
    assert rp.tolist() == [0, 2, 4, 5] and col.tolist() == [1, 1, 0, 0, 0]
    with pytest.raises(ValueError):
        O.Graph(np.array([1, 0], dtype=np.uint32), np.array([0, 1], dtype=np.uint32))


def test_cutoff_closed_form():
    # star K1,4: degrees 4,1,1,1,1 -> mean 8/5, sd sqrt(((2.4)^2 + 4*0.36)/5)
    g = sym_graph(5, [(0, i) for i in range(1, 5)])
    mean = 8 / 5
    sd = math.sqrt(((4 - mean) ** 2 + 4 * (1 - mean) ** 2) / 5)
    assert g.cutoff == math.ceil(mean + 6 * sd)
    assert abs(g.mean - mean) < 1e-12 and abs(g.sd - sd) < 1e-12


def test_peel_and_roots_hand_graph():
    # triangle 0-1-2 with pendant chain 2-3-4 and isolated edge 5-6
    edges = [(0, 1), (1, 2), (0, 2), (2, 3), (3, 4), (5, 6)]
    g = sym_graph(7, edges)
    s, d = directed([(0, 1), (0, 2), (1, 2)])
    pp = O.Prepared(g, O.plan(s, d, 1)).arrays()
    # round 1 removes 4, 5, 6 (degree 1 < 2): 3 > 0.05*7 -> round 2 removes 3
    assert pp["peel_rounds"] == 3
    assert np.diff(pp["rowptr"]).tolist() == [2, 2, 2, 0, 0, 0, 0]
    assert pp["roots"].tolist() == [0, 1, 2]  # cutoff large: boundary 0, no head skip
    assert O.Prepared(g, O.plan(s, d, 1)).count()["count"] == 1
