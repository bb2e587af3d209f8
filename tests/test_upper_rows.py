"""Bitmaps without mirrored bits (Plan.upper_rows) on the CPU: for the plans
the library flags, every row x_q a level intersects with is only used above
x_q, so a neighbourhood bitmap holding just the bits E[i][j], j > i (what a
row's own larger-key range finds, rows in key order) gives the same
enumeration as the full symmetric bitmap.  Checked by brute force on random
neighbourhoods, level by level as the reference enumerates (count_tri /
compute_intersection over one root's bitmap), for every golden and
central-node pattern the library flags; and the flag must be set for the
cliques and clear for plans that read a row below its own position."""
import random

import pytest

import beep_amd as B
from beep_amd import patterns as P


def masks(plan):
    qm = [sum(1 << q for q in plan.qedge[l][1:]) for l in range(plan.width)]
    sm = [sum(1 << s for s in plan.sym[l] if s > 0) for l in range(plan.width)]
    return qm, sm


def walk(l, x, E, d, qm, sm, L):
    if l == L:
        return 1
    n = 0
    for i in range(d):
        if i in x[1:l]:
            continue
        if any((qm[l] >> q) & 1 and not E[x[q]][i] for q in range(1, l)):
            continue
        if any((sm[l] >> s) & 1 and i <= x[s] for s in range(1, l)):
            continue
        x[l] = i
        n += walk(l + 1, x, E, d, qm, sm, L)
    x[l] = None
    return n


def all_plans():
    out = []
    for i, (_, e) in enumerate(P.central207()):
        for mode in (B.SGM, B.SGC):
            out.append((f"central{i}-{mode}", B.Plan.from_edges(*P.directed(e), mode)))
    for name in ("triangle", "4clique", "5clique", "6clique", "diamond", "w5", "tailed_triangle", "4star"):
        for mode in (B.SGM, B.SGC):
            out.append((f"{name}-{mode}", B.Plan.from_edges(*P.directed(P.read_mtx_edges(name)), mode)))
    return out


PLANS = all_plans()
FLAGGED = [(n, p) for n, p in PLANS if p.upper_rows]


def test_flag_on_cliques_only_where_it_holds():
    by = dict(PLANS)
    for name in ("triangle", "4clique", "5clique", "6clique"):
        for mode in (B.SGM, B.SGC):
            assert by[f"{name}-{mode}"].upper_rows, name
    # the diamond's and W5's last levels read rows below their positions
    for name in ("diamond", "w5"):
        assert not by[f"{name}-{B.SGM}"].upper_rows, name
    assert len(FLAGGED) >= 8


@pytest.mark.parametrize("idx", range(len(FLAGGED)))
def test_upper_bitmap_gives_the_same_enumeration(idx):
    name, plan = FLAGGED[idx]
    qm, sm = masks(plan)
    L = plan.width
    rng = random.Random(idx)
    for trial in range(6):
        d = rng.randint(4, 9)
        p = rng.choice((0.4, 0.7, 0.9))
        E = [[False] * d for _ in range(d)]
        for i in range(d):
            for j in range(i + 1, d):
                if rng.random() < p:
                    E[i][j] = E[j][i] = True
        U = [[E[i][j] and j > i for j in range(d)] for i in range(d)]
        for x1 in range(d):
            x = [None] * L
            x[1] = x1
            assert walk(2, list(x), E, d, qm, sm, L) == walk(2, list(x), U, d, qm, sm, L), (name, trial, x1)


def test_a_mutant_rule_fails():
    """Planting the mistake of dropping mirrors for a plan that reads rows
    below their position (the diamond) must change some count: the check
    above is able to fail."""
    plan = B.Plan.from_edges(*P.directed(P.read_mtx_edges("diamond")), B.SGM)
    qm, sm = masks(plan)
    L = plan.width
    rng = random.Random(1)
    differ = False
    for trial in range(20):
        d = 7
        E = [[False] * d for _ in range(d)]
        for i in range(d):
            for j in range(i + 1, d):
                if rng.random() < 0.7:
                    E[i][j] = E[j][i] = True
        U = [[E[i][j] and j > i for j in range(d)] for i in range(d)]
        for x1 in range(d):
            x = [None] * L
            x[1] = x1
            if walk(2, list(x), E, d, qm, sm, L) != walk(2, list(x), U, d, qm, sm, L):
                differ = True
    assert differ
