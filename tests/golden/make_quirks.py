"""Writes tests/golden/quirks.json: graphs on which the reference's root
selection changes the count, with expectations computed here by brute force,
without the C oracle or libbeep.

The root set of count_subgraphs (subgraph_matching.cuh:1052, 1081-1131;
GraphQueue.cuh:85-120; device_utils.cuh:66-94):
  Q0 = {u : max_qdeg <= d(u) < 100 cutoff}  (peeled degrees, cutoff of the
       unpeeled graph, utils.cuh:186-261), ordered by degree descending,
       ties by id ascending;
  b  = #{u in Q0 : d(u) > cutoff};
  R  = Q0 without Q0[0] when the plan is ascending and b >= 1.
Every fixture below makes one part of that rule decide the count:
  window_star   a vertex of degree exactly 100 cutoff (left out) and one of
                100 cutoff - 1 (kept); 800,000 rows so that cutoff = 1;
  skip_head     b >= 1 and Q0[0] (the smallest id among the largest degrees)
                is the smallest key of a 6-clique: the ascending skip drops
                the triangles it roots; a vertex of the same degree with a
                larger id and no larger-key neighbour pins the tie order;
  regular_k6, circulant12
                regular graphs whose degree equals the cutoff: b = 0, so
                nothing is skipped.
For every rule variant a plausible slip would give (<= at the window's upper
bound, >= in the boundary, a skip with b = 0, no skip, descending ties, no
upper bound) the script checks that some expected count changes, so the
fixtures tell the rule from each slip.

Counts follow the constraint system of the enumeration (x_1 >= split when
SYM(1) = [0], adjacency for QEDGE(l)[1:], index order for SYM(l), counting
the first L = k - LUNMAT positions), by brute-force filtering of position
tuples of each root's oriented row.

Run: python tests/golden/make_quirks.py
"""
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from beep_amd.patterns import read_mtx_edges  # noqa: E402
from make_golden import plan_py  # noqa: E402

PATTERNS = ["wedge", "triangle", "tailed_triangle", "diamond", "4clique", "4star"]


def wedge_edges():
    return [(0, 1), (0, 2)]


def pattern_edges(name):
    return wedge_edges() if name == "wedge" else read_mtx_edges(name)


# ------------------------------------------------------------------ graphs
def window_star():
    """Centre A (id 0) with 100 leaves, centre B (id 1) with 99 leaves; the
    leaves of each centre are joined in pairs (so they keep degree 2 through
    the peel of min_qdeg-2 plans; one leaf of B is joined to one of A's
    instead).  The last row id is a leaf so n = 800,000."""
    N = 800_000
    e = set()
    nxt = 2
    la = list(range(nxt, nxt + 100))
    nxt += 100
    lb = list(range(nxt, nxt + 98)) + [N - 1]
    for v in la:
        e.add((0, v))
    for v in lb:
        e.add((1, v))
    for i in range(0, 100, 2):
        e.add((la[i], la[i + 1]))
    for i in range(0, 98, 2):
        e.add((lb[i], lb[i + 1]))
    e.add((la[0], N - 1))  # the odd leaf of B
    return N, sorted(e)


def skip_head():
    """A 6-clique on ids 10..15 (degree 5) and a vertex z = 900 of degree 5
    whose neighbours all have smaller keys; 1,000 rows, so cutoff = 3 and
    b = 7.  Q0[0] is 10, the smallest key of the clique: the ascending skip
    drops the 10 triangles it roots.  Descending ties would put z first,
    whose skip drops nothing."""
    N = 1000
    e = set()
    K = list(range(10, 16))
    for i in range(6):
        for j in range(i + 1, 6):
            e.add((K[i], K[j]))
    z = 900
    for v in (901, 902, 903, 904, 905):
        e.add((z, v))
    e.add((901, 902))
    e.add((903, 904))
    e.add((905, 999))  # 999 has degree 1 (peeled when min_qdeg = 2)
    return N, sorted(e)


def regular_k6():
    return 6, [(a, b) for a in range(6) for b in range(a + 1, 6)]


def circulant12():
    e = set()
    for i in range(12):
        for s in (1, 2):
            j = (i + s) % 12
            e.add((min(i, j), max(i, j)))
    return 12, sorted(e)


GRAPHS = {"window_star": window_star, "skip_head": skip_head, "regular_k6": regular_k6,
          "circulant12": circulant12}


# ------------------------------------------------------------------- prep
RULE = dict(window_le=False, boundary_ge=False, skip_always=False, skip_never=False, tie_desc=False,
            no_upper=False)
SLIPS = ["window_le", "boundary_ge", "skip_always", "skip_never", "tie_desc", "no_upper"]


def prep(n, edges, plan, rule):
    adj = {}
    for a, b in edges:
        adj.setdefault(a, set()).add(b)
        adj.setdefault(b, set()).add(a)
    m = 2 * len(edges)
    mean = m / n
    acc = 0.0
    for v in range(n):  # sequential double sum over every row, as utils.cuh does
        d = len(adj.get(v, ()))
        acc += (d - mean) ** 2
    sd = math.sqrt(acc / n)
    cutoff = math.ceil(mean + 6.0 * sd)
    cur = {v: set(s) for v, s in adj.items()}
    while True:
        S = [v for v in sorted(cur) if 1 <= len(cur[v]) < plan["min_qdeg"]]
        for v in S:
            for w in list(cur[v]):
                cur[w].discard(v)
            cur[v] = set()
        if not len(S) > 0.05 * n:
            break
    deg = {v: len(s) for v, s in cur.items()}
    hi = 100 * cutoff
    q0 = []
    for v in sorted(cur):
        d = deg[v]
        if d < plan["max_qdeg"]:
            continue
        if not rule["no_upper"] and (d > hi if rule["window_le"] else d >= hi):
            continue
        q0.append(v)
    q0.sort(key=lambda v: (-deg[v], -v if rule["tie_desc"] else v))
    bnd = sum(1 for v in q0 if (deg[v] >= cutoff if rule["boundary_ge"] else deg[v] > cutoff))
    skip = plan["asc"] and (bnd >= 1 or rule["skip_always"]) and not rule["skip_never"]
    roots = q0[1:] if skip else q0
    key = (lambda v: (deg[v], v)) if plan["asc"] else (lambda v: (-deg[v], v))
    rows = {u: sorted(cur[u], key=key) for u in roots}
    nlow = {u: sum(1 for w in cur[u] if (deg[w], w) < (deg[u], u)) for u in roots}
    maxd = max(deg.values()) if deg else 0
    head_larger = 0
    if q0:
        h = q0[0]
        head_larger = sum(1 for w in cur[h] if (deg[w], w) > (deg[h], h))
    return dict(cur=cur, rows=rows, roots=roots, nlow=nlow, cutoff=cutoff, bnd=bnd, q0=q0, maxd=maxd,
                head_larger=head_larger)


def count(n, edges, plan, rule):
    pp = prep(n, edges, plan, rule)
    L = plan["k"] - plan["lunmat"]
    res = dict(cutoff=pp["cutoff"], q0_size=len(pp["q0"]), boundary=pp["bnd"], roots=pp["roots"],
               p1=pp["maxd"] < 100 * pp["cutoff"],
               p2=(not plan["asc"]) or pp["bnd"] == 0 or pp["head_larger"] == 0)
    if L <= 2:
        res["count"] = 0
        return res
    skip = plan["sym"][1] == [0]
    qe, sy = plan["qedge"], plan["sym"]
    total = 0
    for u in pp["roots"]:
        row = pp["rows"][u]
        d = len(row)
        nb = pp["cur"]
        x = [0] * L

        def rec(l):
            nonlocal total
            if l == L:
                total += 1
                return
            for i in range(d):
                if i in x[1:l]:
                    continue
                if l == 1:
                    if skip and i < pp["nlow"][u]:
                        continue
                else:
                    v = row[i]
                    if any(v not in nb[row[x[q]]] for q in qe[l][1:]):
                        continue
                    if any(s > 0 and not i > x[s] for s in sy[l]):
                        continue
                x[l] = i
                rec(l + 1)

        rec(1)
    res["count"] = total
    return res


def main():
    plans = {}
    for name in PATTERNS:
        e = pattern_edges(name)
        k = max(max(p) for p in e) + 1
        plans[name] = {"sgm": plan_py(k, e, 0), "sgc": plan_py(k, e, 1)}
    graphs = {g: f() for g, f in GRAPHS.items()}
    cases = []
    bites = {s: [] for s in SLIPS}
    for gname, (n, edges) in graphs.items():
        for pname in PATTERNS:
            for mode in ("sgm", "sgc"):
                plan = plans[pname][mode]
                r = count(n, edges, plan, RULE)
                cases.append(dict(graph=gname, pattern=pname, mode=mode, count=r["count"],
                                  cutoff=r["cutoff"], q0_size=r["q0_size"], boundary=r["boundary"],
                                  n_roots=len(r["roots"]), roots=r["roots"][:64], p1=r["p1"], p2=r["p2"]))
                for s in SLIPS:
                    rule = dict(RULE, **{s: True})
                    c = count(n, edges, plan, rule)["count"]
                    if c != r["count"]:
                        bites[s].append([gname, pname, mode, r["count"], c])
            print(gname, pname, "done", file=sys.stderr, flush=True)
    for s in SLIPS:
        assert bites[s], f"no fixture tells the rule from the slip {s}"
        print(f"slip {s}: changes {len(bites[s])} counts, e.g. {bites[s][0]}", file=sys.stderr)
    out = dict(patterns={p: pattern_edges(p) for p in PATTERNS},
               graphs={g: dict(n=n, edges=e) for g, (n, e) in graphs.items()},
               cases=cases, slips=bites)
    with open(os.path.join(HERE, "quirks.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
