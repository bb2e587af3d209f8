"""beep_count_noalloc -- the launch sequence of beep_count with its scratch
carved from the library's fixed device array instead of allocated -- against
beep_count, through the C ABI: equal counts, tuple-multiset hashes and tuple
buffers on two calls in a row, the second on a smaller input (so it runs over
what the first left in the array), and a clear BEEP_ERANGE for a query whose
scratch does not fit the array."""
import ctypes as C

import numpy as np
import pytest

import beep_amd as B
from beep_amd import _ffi
from beep_amd import patterns as P

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def er_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    a = np.triu(rng.random((n, n)) < p, 1)
    u, v = np.nonzero(a)
    pairs = sorted(set(zip(u.tolist(), v.tolist())) | set(zip(v.tolist(), u.tolist())))
    return B.Graph.from_edges([x for x, _ in pairs], [y for _, y in pairs])


def prepared(g, name, mode):
    s, d = P.directed(P.read_mtx_edges(name))
    pp = B.Prepared(g, B.Plan.from_edges(s, d, mode))
    pp.upload(0)
    return pp


def noalloc_raw(pp, emit, match_cap=0, donate_min=0):
    """The entry point called through ctypes, on the null stream."""
    o = B._opts(0, emit, 0, True, donate_min, 0.0, False, 0)
    buf = None
    if match_cap:
        buf = np.zeros(match_cap * pp.plan.width, dtype=np.uint32)
        o.match_cap = match_cap
        o.match_buf = buf.ctypes.data_as(_ffi.u32p)
    r = _ffi.Result()
    rc = B.lib().beep_count_noalloc(pp._h, pp.plan._h, C.byref(o), None, C.byref(r))
    return rc, r, buf


def rows(m):
    return sorted(map(tuple, np.asarray(m).tolist()))


def test_noalloc_equals_count_two_calls():
    # 1: hubs with bitmaps in the arena and a donation ring (tiny threshold)
    g1 = B.Graph.chung_lu(20000, 60000, 2.2, 5.0, 7)
    for name, mode in (("w5", B.SGM), ("diamond", B.SGM), ("4clique", B.SGM)):
        say("noalloc chunglu20k", name, mode)
        pp = prepared(g1, name, mode)
        want = pp.count(emit=B.EMIT_HASH, donate=True, donate_min=2)
        rc, r, _ = noalloc_raw(pp, B.EMIT_HASH, donate_min=2)
        assert rc == 0, B.lib().beep_last_error()
        assert (r.count, r.hash_sum, r.hash_xor) == (want["count"], want["hash_sum"], want["hash_xor"])
        assert r.stopped == 0
    # 2: a smaller input right after, with the tuple buffer
    g2 = er_graph(60, 0.3, 5)
    for name, mode in (("diamond", B.SGM), ("triangle", B.SGC), ("tailed_triangle", B.SGM)):
        say("noalloc er60", name, mode)
        pp = prepared(g2, name, mode)
        want = pp.count(emit=B.EMIT_BUFFER, match_cap=200000)
        rc, r, buf = noalloc_raw(pp, B.EMIT_BUFFER, match_cap=200000)
        assert rc == 0, B.lib().beep_last_error()
        assert (r.count, r.hash_sum, r.hash_xor) == (want["count"], want["hash_sum"], want["hash_xor"])
        assert r.n_written == want["written"] == want["count"]
        got = buf[:r.n_written * pp.plan.width].reshape(-1, pp.plan.width)
        assert rows(got) == rows(want["matches"])


def test_noalloc_refuses_what_does_not_fit():
    # one vertex adjacent to 30000 others on a cycle: tailed_triangle keeps its
    # full 30000 x 469-word bitmap (112 MB), more than the 64 MiB array
    n = 30000
    src, dst = [], []
    for i in range(1, n + 1):
        j = i % n + 1
        for a, b in ((0, i), (i, 0), (i, j), (j, i)):
            src.append(a)
            dst.append(b)
    order = np.lexsort((np.array(dst), np.array(src)))
    g = B.Graph.from_edges(np.array(src, dtype=np.uint32)[order], np.array(dst, dtype=np.uint32)[order])
    say("noalloc oversize hub")
    pp = prepared(g, "tailed_triangle", B.SGM)
    assert pp.arrays()["n_hub_roots"] >= 1
    rc, _, _ = noalloc_raw(pp, B.EMIT_NONE)
    assert rc == B.ERANGE
    assert "beep_count_noalloc holds at most" in B.lib().beep_last_error().decode()
    # the library stays usable: a query that fits still runs right after
    g2 = er_graph(50, 0.3, 9)
    pp2 = prepared(g2, "4clique", B.SGM)
    want = pp2.count(emit=B.EMIT_HASH)
    rc, r, _ = noalloc_raw(pp2, B.EMIT_HASH)
    assert rc == 0
    assert (r.count, r.hash_sum, r.hash_xor) == (want["count"], want["hash_sum"], want["hash_xor"])
