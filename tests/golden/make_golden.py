"""Writes tests/golden/brute.json: independent expectations for the oracle.

Everything here is computed without the C oracle or libbeep:
  * plans: matching order, back edges, symmetry chain (automorphism orbits by
    trying every permutation), LUNMAT and asc/desc, for the 207 central-node
    patterns and the named queries;
  * counts on tiny graphs, two ways:
      true      #monomorphisms / |Aut(Q)| by exhaustive backtracking;
      semantic  number of maps f (positions -> vertices) satisfying the
                reference's constraint system (root window and head skip,
                x_1 >= split when SYM(1) = [0], adjacency for QEDGE(l)[1:],
                index(f(l)) > index(f(s)) for s in SYM(l), s > 0, counted over
                the first L = k - LUNMAT positions), by brute-force filtering
                of all injective maps -- no bitmaps, no DFS over candidate sets.

Run: python tests/golden/make_golden.py
"""
import itertools
import json
import math
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from beep_amd.patterns import central207, read_mtx_edges, NAMES  # noqa: E402


# ------------------------------------------------------------------ planner
def plan_py(k, edges, sgc):
    adj = [set() for _ in range(k)]
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    deg = [len(adj[v]) for v in range(k)]
    seq, dM, placed = [], [0] * k, [False] * k
    qedge = []
    for i in range(k):
        best, idx = 0, 0
        for j in range(k):
            if not placed[j]:
                lk = dM[j] * k + deg[j]
                if lk > best:
                    best, idx = lk, j
        seq.append(idx)
        placed[idx] = True
        for nb in adj[idx]:
            if not placed[nb]:
                dM[nb] += 1
        qedge.append([j for j in range(i) if seq[j] in adj[idx]])
    padj = [set() for _ in range(k)]
    for i in range(k):
        for q in qedge[i]:
            padj[i].add(q)
            padj[q].add(i)
    pedges = {(a, b) for a in range(k) for b in padj[a]}
    auts = [p for p in itertools.permutations(range(k))
            if all((p[a], p[b]) in pedges for a, b in pedges)]
    fixed, sym = [], [[] for _ in range(k)]
    while True:
        allowed = [p for p in auts if all(p[h] == h for h in fixed)]
        orbit = list(range(k))
        for v in range(k):
            orbit[v] = min(p[v] for p in allowed)  # orbit of v under the stabiliser
        # orbit[v] = smallest image; orbits of a group: min over the orbit
        reps = {}
        for v in range(k):
            reps.setdefault(orbit[v], []).append(v)
        best_rep, best_len = -1, 1
        for r in sorted(reps):
            if len(reps[r]) > best_len:
                best_rep, best_len = r, len(reps[r])
        if best_rep < 0:
            break
        for v in reps[best_rep]:
            if v != best_rep:
                sym[v].append(best_rep)
        fixed.append(best_rep)
    lunmat = 0
    if sgc:
        lunmat = 0 if (k - 2) in qedge[k - 1] else 1
    asc = 1 if sym[1] else 0
    return dict(k=k, seq=seq, qedge=qedge, sym=sym, lunmat=lunmat, asc=asc,
                max_qdeg=max(deg), min_qdeg=min(deg), aut=len(auts))


# -------------------------------------------------------------- graph prep
def prep_py(n, adj, plan):
    """Peel, cutoff, orientation order and roots as the reference defines them."""
    m = sum(len(a) for a in adj)
    degs = [len(a) for a in adj]
    mean = m / n
    sd = math.sqrt(sum((d - mean) ** 2 for d in degs) / n)
    cutoff = math.ceil(mean + 6.0 * sd)
    cur = [set(a) for a in adj]
    while True:
        S = [v for v in range(n) if 1 <= len(cur[v]) < plan["min_qdeg"]]
        for v in S:
            for w in list(cur[v]):
                cur[w].discard(v)
            cur[v] = set()
        if not len(S) > 0.05 * n:
            break
    deg = [len(cur[v]) for v in range(n)]
    if plan["asc"]:
        key = lambda v: (deg[v], v)  # noqa: E731
    else:
        key = lambda v: (-deg[v], v)  # noqa: E731
    rows = {u: sorted(cur[u], key=key) for u in range(n)}
    q0 = [v for v in range(n) if plan["max_qdeg"] <= deg[v] < 100 * cutoff]
    q0.sort(key=lambda v: (-deg[v], v))
    bnd = sum(1 for v in q0 if deg[v] > cutoff)
    roots = q0[1:] if (plan["asc"] and bnd >= 1) else q0
    nlow = {u: sum(1 for w in cur[u] if (deg[w], w) < (deg[u], u)) for u in range(n)}
    return dict(cur=cur, deg=deg, rows=rows, roots=roots, nlow=nlow, cutoff=cutoff, bnd=bnd)


def semantic_count(n, adj, plan):
    """Brute-force count of the reference's constraint system."""
    pp = prep_py(n, adj, plan)
    k, L = plan["k"], plan["k"] - plan["lunmat"]
    if L <= 2:
        return 0
    skip = plan["sym"][1] == [0]
    total = 0
    for u in pp["roots"]:
        row = pp["rows"][u]
        d = len(row)
        nbr = pp["cur"]
        x = [None] * L

        def rec(l):
            # assign position l; every constraint that mentions only positions
            # <= l is checked as soon as position l is set
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
                    if any(row[i] not in nbr[row[x[q]]] for q in plan["qedge"][l][1:]):
                        continue
                    if any(s > 0 and not i > x[s] for s in plan["sym"][l]):
                        continue
                x[l] = i
                rec(l + 1)
                x[l] = None

        rec(1)
    return total


def true_count(n, adj, k, edges, aut):
    pat = [set() for _ in range(k)]
    for a, b in edges:
        pat[a].add(b)
        pat[b].add(a)
    order = sorted(range(k), key=lambda v: -len(pat[v]))
    mono = 0

    def rec(i, f, used):
        nonlocal mono
        if i == k:
            mono += 1
            return
        p = order[i]
        for v in range(n):
            if v in used:
                continue
            if all(v in adj[f[q]] for q in pat[p] if q in f):
                f[p] = v
                used.add(v)
                rec(i + 1, f, used)
                used.discard(v)
                del f[p]

    rec(0, {}, set())
    return mono // aut


def graphs():
    out = []
    rng = random.Random(20261202)
    for n, p in [(8, 0.55), (9, 0.45), (10, 0.4), (10, 0.6), (11, 0.35), (9, 0.7)]:
        e = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]
        out.append((f"er{n}_{int(p * 100)}", n, e))
    out.append(("k7", 7, [(a, b) for a in range(7) for b in range(a + 1, 7)]))
    # wheel with a pendant path: hub 0, rim 1..6, tail 7-8
    w = [(0, i) for i in range(1, 7)] + [(i, i % 6 + 1) for i in range(1, 7)] + [(6, 7), (7, 8)]
    out.append(("wheel_tail", 9, w))
    return out


def main():
    pats = [(f"c{i}", k, e) for i, (k, e) in enumerate(central207())]
    for nm in NAMES:
        e = read_mtx_edges(nm)
        pats.append((nm, max(max(x) for x in e) + 1, e))
    plans = {}
    for name, k, e in pats:
        plans[name] = {"sgm": plan_py(k, e, 0), "sgc": plan_py(k, e, 1)}
    gs = graphs()
    cases = []
    for gname, n, ge in gs:
        adj = [set() for _ in range(n)]
        for a, b in ge:
            adj[a].add(b)
            adj[b].add(a)
        for name, k, e in pats:
            if k > 7:
                continue
            ps, pc = plans[name]["sgm"], plans[name]["sgc"]
            rec = dict(graph=gname, pattern=name, sgm=semantic_count(n, adj, ps),
                       sgc=semantic_count(n, adj, pc), true=true_count(n, adj, k, e, ps["aut"]))
            cases.append(rec)
        print(gname, "done", file=sys.stderr)
    out = dict(patterns={name: dict(k=k, edges=e) for name, k, e in pats}, plans=plans,
               graphs={g: dict(n=n, edges=e) for g, n, e in gs}, cases=cases)
    with open(os.path.join(HERE, "brute.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    exact = sum(1 for c in cases if c["sgm"] == c["true"])
    print(f"{len(cases)} cases; sgm semantics equal the true count in {exact}", file=sys.stderr)


if __name__ == "__main__":
    main()
