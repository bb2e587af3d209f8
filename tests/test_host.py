"""Host logic of libbeep (planner, reader, CSR, cutoff, peel/orient/roots,
generators) against the oracle and the golden data -- no GPU involved."""
import json
import os
import struct

import numpy as np
import pytest

import beep_amd as B
import oracle_py as O
from beep_amd import patterns as P

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "brute.json")))


def directed(edges):
    d = sorted({(a, b) for a, b in edges} | {(b, a) for a, b in edges})
    return [a for a, _ in d], [b for _, b in d]


@pytest.mark.parametrize("name", sorted(GOLD["patterns"]))
def test_product_plan_matches_golden(name):
    s, d = directed(GOLD["patterns"][name]["edges"])
    for mode, key in ((B.SGM, "sgm"), (B.SGC, "sgc")):
        want = GOLD["plans"][name][key]
        p = B.Plan.from_edges(s, d, mode)
        assert p.seq == want["seq"] and p.qedge == want["qedge"] and p.sym == want["sym"]
        assert p.lunmat == want["lunmat"] and p.asc == want["asc"]
        assert (p.max_qdeg, p.min_qdeg) == (want["max_qdeg"], want["min_qdeg"])


def test_plan_from_mtx_files():
    for name in P.NAMES:
        p = B.Plan.from_mtx(P.mtx_path(name), "sgm")
        s, d = P.directed(P.read_mtx_edges(name))
        q = O.plan(np.array(s, dtype=np.uint32), np.array(d, dtype=np.uint32), 0).lists()
        assert p.seq == q["seq"] and p.qedge == q["qedge"] and p.sym == q["sym"]


def test_plan_rejects_bad_queries():
    with pytest.raises(B.BeepError):
        B.Plan.from_edges([0, 1], [1, 0], B.SGM)  # 2 vertices
    with pytest.raises(B.BeepError):
        B.Plan.from_edges([0, 0, 1, 2], [1, 1, 0, 0], B.SGM)  # duplicate edge
    with pytest.raises(ValueError):
        B.Plan.from_edges([0], [1], "tc")


def oracle_graph(g):
    rp, col = g.csr()
    src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    return O.Graph(src, col)


GRAPHS = {
    "chunglu": lambda: B.Graph.chung_lu(6000, 20000, 2.2, 3.0, 5),
    "rmat": lambda: B.Graph.rmat(11, 16 << 11, 0.57, 0.19, 0.19, 3),
    "rmat_cut": lambda: B.Graph.rmat(11, 12 << 11, 0.57, 0.19, 0.19, 4, 1500),
}


@pytest.mark.parametrize("gname", sorted(GRAPHS))
@pytest.mark.parametrize("name", ["triangle", "wedge", "tailed_triangle", "diamond", "4clique", "4star", "w5"])
def test_prepare_matches_oracle(gname, name):
    g = GRAPHS[gname]()
    og = oracle_graph(g)
    assert (g.n, g.m, g.cutoff) == (og.n, og.m, og.cutoff)
    s, d = P.directed(P.read_mtx_edges(name))
    for mode in (B.SGM, B.SGC):
        a = B.Prepared(g, B.Plan.from_edges(s, d, mode)).arrays()
        b = O.Prepared(og, O.plan(np.array(s, dtype=np.uint32), np.array(d, dtype=np.uint32), mode)).arrays()
        for k in ("rowptr", "col", "ocol", "nlow", "roots"):
            assert np.array_equal(a[k], b[k]), k
        for k in ("q0_size", "boundary", "peel_rounds", "asc", "max_deg"):
            assert a[k] == b[k], k


def test_alg_bytes_matches_oracle():
    g = GRAPHS["chunglu"]()
    og = oracle_graph(g)
    s, d = P.directed(P.read_mtx_edges("diamond"))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, B.SGM))
    op = O.Prepared(og, O.plan(np.array(s, dtype=np.uint32), np.array(d, dtype=np.uint32), 0))
    assert pp.alg_bytes == op.alg_bytes()


def test_generators_deterministic():
    a = B.Graph.rmat(12, 40000, 0.57, 0.19, 0.19, 7)
    b = B.Graph.rmat(12, 40000, 0.57, 0.19, 0.19, 7)
    c = B.Graph.rmat(12, 40000, 0.57, 0.19, 0.19, 8)
    ra, ca = a.csr()
    rb, cb = b.csr()
    assert np.array_equal(ra, rb) and np.array_equal(ca, cb)
    assert not np.array_equal(ca, c.csr()[1][:len(ca)]) or len(ca) != c.m
    # symmetric, no loops, no duplicates, sorted rows
    src = np.repeat(np.arange(a.n), np.diff(ra).astype(np.int64))
    assert not np.any(src == ca)
    pairs = set(zip(src.tolist(), ca.tolist()))
    assert len(pairs) == len(ca) and all((y, x) in pairs for x, y in pairs)
    for v in range(a.n):
        row = ca[ra[v]:ra[v + 1]]
        assert np.all(row[1:] > row[:-1])


def test_rmat_rejection_bound():
    g = B.Graph.rmat(11, 20000, 0.57, 0.19, 0.19, 1, 1000)
    rp, col = g.csr()
    assert g.n <= 1000 and (len(col) == 0 or col.max() < 1000)


def test_bel_roundtrip_and_readers_agree(tmp_path):
    g = GRAPHS["chunglu"]()
    path = str(tmp_path / "g.bel")
    g.write_bel(path)
    h = B.Graph.load(path)
    assert np.array_equal(g.csr()[0], h.csr()[0]) and np.array_equal(g.csr()[1], h.csr()[1])
    s, d = O.read_edges(path)
    rp, col = g.csr()
    assert np.array_equal(d, col)
    raw = open(path, "rb").read(48)
    dst0, src0, w0 = struct.unpack("<QQQ", raw[:24])
    assert (src0, dst0, w0) == (int(s[0]), int(d[0]), 0)


def test_mtx_and_tsv_readers_agree_with_oracle(tmp_path):
    mtx = tmp_path / "q.mtx"
    mtx.write_text("%%MatrixMarket matrix coordinate real general\n4 4 6\n0 1 1.5\n0 2\n1 0\n\n2 0 3\n")
    tsv = tmp_path / "q.tsv"
    tsv.write_text("1 0 1\n2 0 1\n0 1 1\n0 2 1\n")
    for p in (mtx, tsv):
        s, d = O.read_edges(str(p))
        g = B.Graph.load(str(p))
        og = O.Graph(s, d)
        assert np.array_equal(g.csr()[0], og.arrays()[0]) and np.array_equal(g.csr()[1], og.arrays()[1])


def test_sort_edges_flag(tmp_path):
    path = tmp_path / "u.bel"
    recs = [(2, 0), (1, 0), (0, 1), (0, 2)]  # (dst, src), row 0 unsorted
    path.write_bytes(b"".join(struct.pack("<QQQ", a, b, 0) for a, b in recs))
    g = B.Graph.load(str(path), sort_edges=True)
    assert g.csr()[1].tolist() == [1, 2, 0, 0]
    g2 = B.Graph.load(str(path))
    assert g2.csr()[1].tolist() == [2, 1, 0, 0]


def test_load_errors():
    with pytest.raises(B.BeepError) as e:
        B.Graph.load("/nonexistent/x.bel")
    assert e.value.code == -5
    with pytest.raises(B.BeepError) as e:
        B.Graph.load("/tmp/x.unknown")
    assert e.value.code == -22
    with pytest.raises(B.BeepError):
        B.Graph.from_edges([1, 0], [0, 1])  # not grouped by ascending source


def test_unsorted_rows_are_refused():
    """Rows out of id order (a .bel read without -s) are refused by both the
    product and the oracle: the reference's bitmap build binary-searches the
    id-ordered row (host_utils.cuh:284-316), so its counts are defined only
    for sorted rows.  The same edges sorted are accepted."""
    import numpy as np
    import oracle_py as O
    tri = [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3)]
    e = sorted({(a, b) for a, b in tri} | {(b, a) for a, b in tri})
    src = [a for a, _ in e]
    dst = [b for _, b in e]
    # row 1 = [0, 2, 3] written as [3, 0, 2]
    i = src.index(1)
    dst[i:i + 3] = [3, 0, 2]
    g = B.Graph.from_edges(src, dst)
    plan = B.Plan.from_undirected([(0, 1), (1, 2), (0, 2)], "sgc")
    with pytest.raises(B.BeepError) as ei:
        B.Prepared(g, plan)
    assert "not sorted" in str(ei.value)
    s, d = P.directed([(0, 1), (1, 2), (0, 2)])
    with pytest.raises(ValueError):
        O.Prepared(O.Graph(np.array(src, dtype=np.uint32), np.array(dst, dtype=np.uint32)), O.plan(s, d, 1))
    gs = B.Graph.from_edges([a for a, _ in e], [b for _, b in e])
    assert B.Prepared(gs, plan).arrays()["q0_size"] == 4


def test_prep_subset_keeps_ranks_and_counts_add_up():
    """beep_prep_subset takes the roots at the given ranks of the parent's
    root list; the oracle's count over ranks adds up to its full count."""
    g = B.Graph.chung_lu(3000, 9000, 2.2, 3.0, 5)
    og = oracle_graph(g)
    for name, mode in (("diamond", B.SGM), ("4clique", B.SGC), ("w5", B.SGM)):
        s, d = directed(P.read_mtx_edges(name))
        pp = B.Prepared(g, B.Plan.from_edges(s, d, mode))
        roots = pp.arrays()["roots"]
        ranks = np.arange(1, len(roots), 3, dtype=np.uint64)
        sub = pp.subset(ranks)
        assert sub.arrays()["roots"].tolist() == roots[ranks.astype(np.int64)].tolist()
        op = O.Prepared(og, O.plan(s, d, mode))
        assert op.arrays()["roots"].tolist() == roots.tolist()
        parts = [np.arange(k, len(roots), 3, dtype=np.uint64) for k in range(3)]
        total = sum(op.count_ranks(r)["count"] for r in parts)
        assert total == op.count()["count"]
    with pytest.raises(B.BeepError):
        pp.subset([2, 1])
    with pytest.raises(B.BeepError):
        pp.subset([len(roots)])
