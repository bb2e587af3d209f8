"""The paired last levels (KPlan::pair) on the CPU: for plans whose last three
levels a = L-3, b = L-2, c = L-1 satisfy
    qmask[c] = qmask[a] + {b},  smask[c] = smask[a] + {a},
    a in qmask[b], qmask[b] - {a} and smask[b] below a,
the kernels replace the walk over x_a, x_b, x_c by
    sum over x_b in B of C(|S n E[x_b]|, 2)
(S = C_a, B = the candidates of level b without its constraints on x_a).
This checks that identity by brute force: the level-by-level enumeration of
the reference (count_tri / compute_intersection semantics over one root's
neighbourhood bitmap) against the paired sum, on random neighbourhoods, for
every central-node pattern of 5-7 vertices whose plan qualifies."""
import random

import pytest

import beep_amd as B
from beep_amd import patterns as P


def masks(plan):
    qm = [sum(1 << q for q in plan.qedge[l][1:]) for l in range(plan.width)]
    sm = [sum(1 << s for s in plan.sym[l] if s > 0) for l in range(plan.width)]
    return qm, sm


def qualifies(plan):
    L = plan.width
    if L < 5:
        return False
    qm, sm = masks(plan)
    a, b, c = L - 3, L - 2, L - 1
    below = (1 << a) - 1
    return (qm[c] == qm[a] | (1 << b) and (qm[b] >> a) & 1 and not qm[b] & ~(1 << a) & ~below
            and sm[c] == sm[a] | (1 << a) and not sm[b] & ~below)


PAIRED = [i for i, (_, e) in enumerate(P.central207())
          if qualifies(B.Plan.from_edges(*P.directed(e), B.SGM))]


def candidates(l, x, E, d, qm, sm, skip=0):
    """C_l given x_1..x_{l-1}: not yet chosen, adjacent to x_q for q in qmask[l]
    (minus the positions in `skip`), above x_s for s in smask[l]."""
    out = []
    for i in range(d):
        if i in x[1:l]:
            continue
        if any((qm[l] >> q) & 1 and not (skip >> q) & 1 and not E[x[q]][i] for q in range(1, l)):
            continue
        if any((sm[l] >> s) & 1 and not (skip >> s) & 1 and i <= x[s] for s in range(1, l)):
            continue
        out.append(i)
    return out


def walk(l, x, E, d, qm, sm, L):
    if l == L:
        return 1
    n = 0
    for i in candidates(l, x, E, d, qm, sm):
        x[l] = i
        n += walk(l + 1, x, E, d, qm, sm, L)
    x[l] = None
    return n


def paired(l, x, E, d, qm, sm, L):
    a = L - 3
    if l == a:
        S = set(candidates(a, x, E, d, qm, sm))
        x_b = candidates(a + 1, x + [None], E, d, qm, sm, skip=1 << a)
        return sum(len([i for i in S if E[xb][i]]) * (len([i for i in S if E[xb][i]]) - 1) // 2 for xb in x_b)
    n = 0
    for i in candidates(l, x, E, d, qm, sm):
        x[l] = i
        n += paired(l + 1, x, E, d, qm, sm, L)
    x[l] = None
    return n


def test_some_plans_qualify():
    # W5 (hub + 4-cycle) is one of them
    w5 = B.Plan.from_edges(*P.directed(P.read_mtx_edges("w5")), B.SGM)
    assert qualifies(w5)
    assert len(PAIRED) >= 2


@pytest.mark.parametrize("idx", PAIRED)
def test_paired_levels_equal_the_walk(idx):
    plan = B.Plan.from_edges(*P.directed(P.central207()[idx][1]), B.SGM)
    qm, sm = masks(plan)
    L = plan.width
    rng = random.Random(idx)
    for trial in range(12):
        d = rng.randint(5, 11)
        p = rng.choice((0.3, 0.5, 0.8))
        E = [[False] * d for _ in range(d)]
        for i in range(d):
            for j in range(i + 1, d):
                if rng.random() < p:
                    E[i][j] = E[j][i] = True
        for x1 in range(d):
            x = [None] * L
            x[1] = x1
            assert walk(2, x, E, d, qm, sm, L) == paired(2, list(x), E, d, qm, sm, L), (idx, trial, x1)
