"""Root selection where it decides the count (tests/golden/quirks.json, written
by make_quirks.py without the oracle or libbeep): the root window's upper
bound 100 cutoff, the boundary b = #{d > cutoff} and the ascending head skip
(subgraph_matching.cuh:1052, 1081-1131; device_utils.cuh:66-94).

CPU only: the oracle's counts and root sets, and the product's root sets and
preconditions (beep_prepare runs on the host), equal the brute-force
expectations.  The GPU counts are checked in test_gpu_parity.py."""
import json
import os

import numpy as np
import pytest

import beep_amd as B
import oracle_py as O
from beep_amd import patterns as P

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "quirks.json")))
_G = {}


def directed_edges(edges):
    e = sorted({(a, b) for a, b in edges} | {(b, a) for a, b in edges})
    return np.array([a for a, _ in e], dtype=np.uint32), np.array([b for _, b in e], dtype=np.uint32)


def graphs(name):
    if name not in _G:
        g = GOLD["graphs"][name]
        s, d = directed_edges(g["edges"])
        _G[name] = (B.Graph.from_edges(s, d), O.Graph(s, d))
        assert _G[name][0].n == g["n"]
    return _G[name]


def ids(c):
    return f"{c['graph']}-{c['pattern']}-{c['mode']}"


def test_every_slip_is_told_apart():
    """Each plausible mistake in the rule changes some expected count."""
    for slip in ("window_le", "boundary_ge", "skip_always", "skip_never", "tie_desc", "no_upper"):
        assert GOLD["slips"][slip], slip


def test_fixtures_make_the_quirks_bite():
    by = {(c["graph"], c["pattern"], c["mode"]): c for c in GOLD["cases"]}
    assert not by[("window_star", "4star", "sgm")]["p1"]          # a vertex above the window
    assert by[("window_star", "4star", "sgm")]["cutoff"] == 1
    assert not by[("skip_head", "triangle", "sgm")]["p2"]         # the skip drops triangles
    assert by[("skip_head", "triangle", "sgm")]["boundary"] >= 1
    assert by[("regular_k6", "triangle", "sgm")]["boundary"] == 0  # degree == cutoff: no skip


@pytest.mark.parametrize("c", GOLD["cases"], ids=ids)
def test_oracle_matches_brute_force(c):
    bg, og = graphs(c["graph"])
    s, d = P.directed(GOLD["patterns"][c["pattern"]])
    op = O.Prepared(og, O.plan(s, d, B.mode_of(c["mode"])))
    a = op.arrays()
    assert (a["q0_size"], a["boundary"], a["cutoff"]) == (c["q0_size"], c["boundary"], c["cutoff"])
    assert len(a["roots"]) == c["n_roots"]
    assert a["roots"][:64].tolist() == c["roots"]
    assert op.count(want_hash=False)["count"] == c["count"]


@pytest.mark.parametrize("c", GOLD["cases"], ids=ids)
def test_product_root_set_and_preconditions(c):
    bg, og = graphs(c["graph"])
    s, d = P.directed(GOLD["patterns"][c["pattern"]])
    pp = B.Prepared(bg, B.Plan.from_edges(s, d, c["mode"]))
    a = pp.arrays()
    assert (a["q0_size"], a["boundary"], a["cutoff"]) == (c["q0_size"], c["boundary"], c["cutoff"])
    assert len(a["roots"]) == c["n_roots"]
    assert a["roots"][:64].tolist() == c["roots"]
    assert (a["p1"], a["p2"]) == (c["p1"], c["p2"])
