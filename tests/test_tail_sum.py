"""The summed last level (KPlan::tail) on the CPU: the kernels' closed forms
for sum over x in S of |F n (x, inf)| (or |F minus {x}|), evaluated per 64-bit
word with prefix sums of |F_w| as the hub waves (rounds of 64 words with a
carry) and the LDS groups (one word per lane) evaluate them, against brute
force over the bits.  Pins the formulas; the kernels themselves are pinned
by the count-only GPU tests against the oracle."""
import random

M = (1 << 64) - 1


def popc(x):
    return bin(x).count("1")


def pairs_in_word(S, F):
    if S == F:
        k = popc(S)
        return k * (k - 1) // 2
    n = 0
    if popc(S) <= popc(F):
        while S:
            b = (S & -S).bit_length() - 1
            S &= S - 1
            n += popc((F >> b) >> 1)
    else:
        while F:
            b = (F & -F).bit_length() - 1
            F &= F - 1
            n += popc(S & ((1 << b) - 1))
    return n


def lt_mask(x, w):
    if (x >> 6) > w:
        return M
    if (x >> 6) < w:
        return 0
    return (1 << (x & 63)) - 1


def brute(Sw, Fw, gt):
    S = [w * 64 + b for w, x in enumerate(Sw) for b in range(64) if (x >> b) & 1]
    F = [w * 64 + b for w, x in enumerate(Fw) for b in range(64) if (x >> b) & 1]
    if gt:
        return sum(sum(1 for y in F if y > x) for x in S)
    fs = set(F)
    return sum(len(F) - (x in fs) for x in S)


def hub_rounds(Sw, Fw, ws, fl, pos0, hi2, gt):
    """hub_dfs_kernel's tail_count: S from word ws and bits [pos0, hi2), F from word fl."""
    W = len(Sw)
    w0 = ws if gt else min(ws, fl)
    acc, sl, carry = 0, 0, 0
    for rr in range(w0 >> 6, (W + 63) // 64):
        words = []
        for lane in range(64):
            w = rr * 64 + lane
            S = F = 0
            if w < W and w >= w0:
                if w >= ws:
                    S = Sw[w] & (~lt_mask(pos0, w) & M) & lt_mask(hi2, w)
                if w >= fl:
                    F = Fw[w]
            words.append((S, F))
        run = 0
        for S, F in words:
            run += popc(F)
            fin = run + carry
            acc += popc(S) * fin - pairs_in_word(S, F) if gt else popc(S & F)
            sl += popc(S)
        carry += run
    return (sl * carry - acc) & M


def lds_group(Sw, Fw, gt):
    """lds_batch_kernel's tail_sum: one word per lane of a P-lane group."""
    P = 1
    while P < len(Sw):
        P *= 2
    pre, run = [], 0
    for lx in range(P):
        run += popc(Fw[lx]) if lx < len(Fw) else 0
        pre.append(run)
    tot = 0
    for lx in range(P):
        S = Sw[lx] if lx < len(Sw) else 0
        F = Fw[lx] if lx < len(Fw) else 0
        tot += popc(S) * (run - pre[lx]) + pairs_in_word(S, F) if gt else popc(S) * run - popc(S & F)
    return tot


def rand_words(rng, W, d, dens, lo=0):
    return [0 if w < lo else sum(1 << b for b in range(64) if w * 64 + b < d and rng.random() < dens)
            for w in range(W)]


def test_hub_rounds_match_brute_force():
    rng = random.Random(5)
    for _ in range(80):
        W = rng.randint(1, 140)
        d = rng.randint((W - 1) * 64 + 1, W * 64)
        dens = rng.choice([0.02, 0.2, 0.6, 1.0])
        fl = rng.randint(0, W - 1) if rng.random() < 0.5 else 0
        Sw = rand_words(rng, W, d, dens)
        Fw = Sw[:] if rng.random() < 0.3 else rand_words(rng, W, d, dens, fl)
        if rng.random() < 0.3:
            Fw = [x if w >= fl else 0 for w, x in enumerate(Fw)]
        wl = rng.randint(0, W - 1) if rng.random() < 0.3 else 0
        pos0 = rng.randint(0, d) if rng.random() < 0.5 else 0
        hi2 = rng.randint(pos0, d + 5) if rng.random() < 0.5 else 0xFFFFFFFF
        gt = rng.random() < 0.5
        ws = max(pos0 >> 6, wl)
        Sr = [(x & (~lt_mask(pos0, w) & M) & lt_mask(hi2, w)) if w >= ws else 0 for w, x in enumerate(Sw)]
        fl_eff = fl if all(x == 0 for x in Fw[:fl]) else 0
        assert hub_rounds(Sw, Fw, ws, fl_eff, pos0, hi2, gt) == brute(Sr, Fw, gt)


def test_lds_groups_match_brute_force():
    rng = random.Random(6)
    for _ in range(300):
        W = rng.randint(1, 64)
        d = rng.randint((W - 1) * 64 + 1, W * 64)
        dens = rng.choice([0.05, 0.3, 0.8])
        Sw = rand_words(rng, W, d, dens)
        Fw = Sw[:] if rng.random() < 0.3 else rand_words(rng, W, d, dens)
        gt = rng.random() < 0.5
        assert lds_group(Sw, Fw, gt) == brute(Sw, Fw, gt)


def test_pairs_in_word_cases():
    assert pairs_in_word(0b1011, 0b1011) == 3
    assert pairs_in_word(0b0001, 0b1110) == 3
    assert pairs_in_word(0b1110, 0b0001) == 0
    assert pairs_in_word(1 << 63, M) == 0
    assert pairs_in_word(1, M) == 63
