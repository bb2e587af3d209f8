"""The paired last levels of the hub DFS (KPlan::pair) on the CPU: the
kernel's bit-sliced evaluation of

    sum over x_b in B of C(|S n E[x_b]|, 2)

-- S's members listed, every round of 64 words keeping c_p = |{s in S :
E[s][p]}| of each position p as slices (slice j = bit j of every c_p),
member words added with a ripple carry, and the slices turned into
sum_p C(c_p, 2) at the round's end -- against brute force on random
symmetric neighbourhood bitmaps.  A planted asymmetric bitmap (one bit of a
pair left out, as the upper-row builds do) makes the two differ, which is
why the kernel takes this form for pair plans only (they keep both bits).
The kernels themselves are pinned by the count-only GPU tests (W5 and the
other pair plans) against the oracle."""
import random

SLICES = 20


def popc(x):
    return bin(x).count("1")


def sliced_pair_sum(E, S, B, W):
    """E: list of rows (lists of W 64-bit words); S, B: lists of W words."""
    members = [(w << 6) + b for w in range(W) for b in range(64) if (S[w] >> b) & 1]
    assert len(members) < (1 << SLICES)
    nb = max(len(members), 1).bit_length()
    total = 0
    for w in range(W):  # one lane's word of one round
        bc = B[w]
        if not bc:
            continue
        sl = [0] * SLICES
        for s in members:
            c = E[s][w] & bc
            for j in range(SLICES):  # ripple carry
                if c:
                    t = sl[j] & c
                    sl[j] ^= c
                    c = t
        acc = 0
        for j in range(1, nb):
            acc += (1 << (j - 1)) * ((1 << j) - 1) * popc(sl[j])
            for i in range(j):
                acc += popc(sl[i] & sl[j]) << (i + j)
        total += acc
    return total


def brute(E, S, B, d):
    Sset = {i for i in range(d) if (S[i >> 6] >> (i & 63)) & 1}
    n = 0
    for xb in range(d):
        if (B[xb >> 6] >> (xb & 63)) & 1:
            k = sum(1 for s in Sset if (E[xb][s >> 6] >> (s & 63)) & 1)
            n += k * (k - 1) // 2
    return n


def random_case(rng, d, p, ps, pb):
    W = (d + 63) // 64
    E = [[0] * W for _ in range(d)]
    for i in range(d):
        for j in range(i + 1, d):
            if rng.random() < p:
                E[i][j >> 6] |= 1 << (j & 63)
                E[j][i >> 6] |= 1 << (i & 63)
    S = [0] * W
    B = [0] * W
    for i in range(d):
        if rng.random() < ps:
            S[i >> 6] |= 1 << (i & 63)
        if rng.random() < pb:
            B[i >> 6] |= 1 << (i & 63)
    return E, S, B, W


def test_sliced_pair_sum_matches_brute_force():
    rng = random.Random(7)
    for d, p, ps, pb in [(5, 0.6, 0.8, 0.9), (64, 0.3, 0.4, 0.7), (130, 0.5, 0.3, 0.5),
                         (200, 0.9, 0.9, 1.0), (97, 0.05, 0.5, 0.5), (150, 0.2, 0.02, 0.9)]:
        E, S, B, W = random_case(rng, d, p, ps, pb)
        assert sliced_pair_sum(E, S, B, W) == brute(E, S, B, d), (d, p, ps, pb)


def test_counts_need_the_high_slices():
    # every member of S adjacent to one x_b: c_p = |S| > 2^k for several k
    d = 300
    W = (d + 63) // 64
    E = [[0] * W for _ in range(d)]
    hub = 299
    for s in range(0, 290):
        E[s][hub >> 6] |= 1 << (hub & 63)
        E[hub][s >> 6] |= 1 << (s & 63)
    S = [0] * W
    for s in range(290):
        S[s >> 6] |= 1 << (s & 63)
    B = [0] * W
    B[hub >> 6] |= 1 << (hub & 63)
    assert sliced_pair_sum(E, S, B, W) == 290 * 289 // 2 == brute(E, S, B, d)


def test_asymmetric_bitmap_differs():
    rng = random.Random(11)
    E, S, B, W = random_case(rng, 90, 0.4, 0.5, 0.8)
    want = sliced_pair_sum(E, S, B, W)
    assert want == brute(E, S, B, 90)
    # keep only the bit in the row of the smaller index for one pair between a
    # member of S and a candidate of B
    for s in range(90):
        if not (S[s >> 6] >> (s & 63)) & 1:
            continue
        for xb in range(s + 1, 90):
            if (B[xb >> 6] >> (xb & 63)) & 1 and (E[xb][s >> 6] >> (s & 63)) & 1:
                E[xb][s >> 6] &= ~(1 << (s & 63))
                assert sliced_pair_sum(E, S, B, W) != brute(E, S, B, 90)
                return
    raise AssertionError("no pair to plant")


def donated_word_ranges(W, r0, c2, word_lo, chunk_work=1 << 12):
    """The hub kernel's cut of a 5-vertex pair item into donated parts (whole
    rounds of 64 words, about chunk_work member-rounds each; the first part
    starts at position 0) and the word range [blo, bhi) each part's paired
    pass then covers, with B's own first word word_lo."""
    R = (W + 63) // 64
    nr = R - r0 if R > r0 else 0
    rpc = max(1, chunk_work // c2 if c2 else nr)
    nc = (nr + rpc - 1) // rpc
    out = []
    for c in range(nc):
        lo = (r0 + c * rpc) << 12 if c else 0
        hi = (W << 6) if c + 1 == nc else (r0 + (c + 1) * rpc) << 12
        blo = max(word_lo, lo >> 6)
        bhi = W if hi >= (W << 6) else hi >> 6
        out.append((blo, bhi))
    return out


def test_donated_parts_cover_b_once():
    rng = random.Random(3)
    for _ in range(2000):
        W = rng.randint(65, 20000)
        R = (W + 63) // 64
        word_lo = rng.randint(0, W - 1)
        r0 = word_lo >> 6
        c2 = rng.choice([1, 2, 7, 63, 64, 65, 1000, 4096, 5000, 70000])
        parts = donated_word_ranges(W, r0, c2, word_lo)
        covered = []
        for blo, bhi in parts:
            covered.extend(range(blo, max(blo, bhi)))
        assert covered == list(range(word_lo, W)), (W, word_lo, c2)
        # parts after the first start on a round (64 words)
        assert all(blo % 64 == 0 for blo, _ in parts[1:]), (W, word_lo, c2)
        assert len(parts) <= R - r0
