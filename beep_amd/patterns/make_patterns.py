"""Writes the pattern library shipped with beep_amd.

* named .mtx queries (0-based, both directions listed, as README.md:51-66 of
  the reference describes): triangle, wedge, tailed_triangle, diamond, 4clique,
  4star, w5 (hub + 4-cycle), 5clique, 6clique;
* central207.txt: every graph on 2..6 vertices up to isomorphism (2 + 4 + 11 +
  34 + 156 = 207), each coned with a universal vertex 0, i.e. all central-node
  patterns on 3..7 vertices.  One line per pattern: "k e a-b a-b ...".

Run: python beep_amd/patterns/make_patterns.py
"""
import itertools
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

NAMED = {
    "triangle": [(0, 1), (0, 2), (1, 2)],
    "wedge": [(0, 1), (0, 2)],
    "tailed_triangle": [(0, 1), (0, 2), (0, 3), (1, 2)],
    "diamond": [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)],
    "4clique": [(a, b) for a in range(4) for b in range(a + 1, 4)],
    "4star": [(0, 1), (0, 2), (0, 3)],
    "w5": [(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (2, 3), (3, 4), (4, 1)],
    "5clique": [(a, b) for a in range(5) for b in range(a + 1, 5)],
    "6clique": [(a, b) for a in range(6) for b in range(a + 1, 6)],
}


def mtx_text(edges):
    k = max(max(e) for e in edges) + 1
    d = sorted({(a, b) for a, b in edges} | {(b, a) for a, b in edges})
    lines = [f"{k} {k} {len(d)}"] + [f"{a} {b}" for a, b in d]
    return "\n".join(lines) + "\n"


def graphs_up_to_iso(n):
    """Canonical representatives of all graphs on n vertices (min edge bitmask)."""
    pairs = list(itertools.combinations(range(n), 2))
    npairs = len(pairs)
    idx = {p: i for i, p in enumerate(pairs)}
    masks = np.arange(1 << npairs, dtype=np.int64)
    bits = ((masks[:, None] >> np.arange(npairs)) & 1).astype(np.int64)
    canon = masks.copy()
    for perm in itertools.permutations(range(n)):
        target = np.array([idx[tuple(sorted((perm[a], perm[b])))] for a, b in pairs])
        permuted = (bits << target[None, :]).sum(axis=1)
        np.minimum(canon, permuted, out=canon)
    reps = sorted(set(canon.tolist()))
    return [[pairs[i] for i in range(npairs) if (r >> i) & 1] for r in reps]


def central_patterns():
    out = []
    for n in range(2, 7):
        for g in graphs_up_to_iso(n):
            edges = [(0, v + 1) for v in range(n)] + [(a + 1, b + 1) for a, b in g]
            out.append((n + 1, sorted(edges)))
    return out


def main():
    for name, e in NAMED.items():
        with open(os.path.join(HERE, f"{name}.mtx"), "w") as f:
            f.write(mtx_text(e))
    pats = central_patterns()
    assert len(pats) == 207, len(pats)
    with open(os.path.join(HERE, "central207.txt"), "w") as f:
        for k, e in pats:
            f.write(f"{k} {len(e)} " + " ".join(f"{a}-{b}" for a, b in e) + "\n")
    print("wrote", len(NAMED), "named queries and", len(pats), "central-node patterns")


if __name__ == "__main__":
    main()
