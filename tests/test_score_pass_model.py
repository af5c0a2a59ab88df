"""CPU model of the int16 score pass's arithmetic, checked against the
reference recurrence (SmithWatermanGotoh.cs:154-214) and the oracle.

The kernel keeps the horizontal gap state as H' = max(H, 0), updated with an
unsigned saturating subtract: max(H' - e, 0) = max(H - e, 0), so
H' = max(max(H' - e, 0), V_left - o) = max(H, 0) and the cell value
V = max(f, G, H, 0) = max(f, G, H') needs no separate 0 floor.  It tracks,
per row, the running maximum and the first 16-column block in which it was
reached, folding each block's maximum in at the block's end.  This model
runs both forms on random pairs and requires the same V everywhere and the
oracle's first-maximum cell inside the reported block."""
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle import random_related

BLOCK = 16  # kFirstMaxBlock
NEG = -(1 << 20)


def dp_rows(s1, s2, M, o, e, clamped):
    """Yields V rows 1..m of the local DP; `clamped` uses the kernel's H'."""
    n = len(s2)
    Vprev = np.zeros(n + 1, dtype=np.int64)
    Gprev = np.full(n + 1, NEG, dtype=np.int64)
    for i in range(1, len(s1) + 1):
        V = np.zeros(n + 1, dtype=np.int64)
        G = np.full(n + 1, NEG, dtype=np.int64)
        H = 0 if clamped else NEG
        for j in range(1, n + 1):
            f = Vprev[j - 1] + M[s1[i - 1], s2[j - 1]]
            G[j] = max(Gprev[j] - e, Vprev[j] - o)
            if clamped:
                H = max(max(H - e, 0), V[j - 1] - o)
                V[j] = max(f, G[j], H)
            else:
                H = max(H - e, V[j - 1] - o)
                V[j] = max(f, G[j], H, 0)
        yield V
        Vprev, Gprev = V, G


def first_max_block(s1, s2, M, o, e):
    """The kernel's tracking: per row the running maximum rb and the block of
    its first occurrence, block maxima folded in at block ends; rows in
    order, strict improvement."""
    best, bi, jb = -1, 0, 0
    for i, V in enumerate(dp_rows(s1, s2, M, o, e, True), start=1):
        rb, rj = -1, 0
        for b0 in range(1, len(s2) + 1, BLOCK):
            bm = int(V[b0:b0 + BLOCK].max())
            if bm > rb:
                rb, rj = bm, b0
        if rb > best:
            best, bi, jb = rb, i, rj
    return best, bi, jb


def cases(ednafull, blosum62):
    rng = random.Random(31)
    out = []
    for _ in range(60):
        a, b = random_related(rng, rng.randint(1, 70))
        out.append((a, b, ednafull, rng.choice([16, 10, 5]), rng.choice([4, 1, 2])))
    aa = b"ARNDCQEGHILKMFPSTWYVX"
    for _ in range(30):
        a = bytes(rng.choice(aa) for _ in range(rng.randint(1, 60)))
        b = bytes(rng.choice(aa) for _ in range(rng.randint(1, 60)))
        out.append((a, b, blosum62, 10, 1))
    out.append((b"A" * 40, b"A" * 37, ednafull, 16, 4))
    out.append((b"ACGT" * 12, b"TGCA" * 9, ednafull, 3, 3))
    return out


def test_clamped_horizontal_state_gives_the_same_values(ednafull, blosum62):
    for a, b, M, o, e in cases(ednafull, blosum62):
        Mi = M.astype(np.int64)
        for V0, V1 in zip(dp_rows(a, b, Mi, o, e, False), dp_rows(a, b, Mi, o, e, True)):
            assert np.array_equal(V0, V1), (a, b, o, e)


def test_first_maximum_lies_in_the_tracked_block(ednafull, blosum62):
    for a, b, M, o, e in cases(ednafull, blosum62):
        best, bi, jb = first_max_block(a, b, M.astype(np.int64), o, e)
        r = O.sw_literal(a, b, M, o, e)
        assert best == int(r["score"]), (a, b)
        assert bi == r["best_row"], (a, b, bi, r["best_row"])
        assert jb <= r["best_col"] < jb + BLOCK, (a, b, jb, r["best_col"])
