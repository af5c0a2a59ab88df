"""The two-pass schedule's re-run rectangle, checked on the CPU.

The GPU re-run recomputes the statistics of the reference's traceback
(SmithWatermanGotoh.Traceback, SmithWatermanGotoh.cs:229-332) only over the
rectangle [si..bi] x [sj..wend]: (bi, jb) = the first-maximum row and the
first column of its 4-column block, wend = min(jb + 3, |s2|), and (si, sj) =
the topmost row / leftmost column where a reverse DP anchored on the block
reaches the first-maximum score.  This restates that reverse DP in Python and
checks, against the oracle's literal aligner on the full pair, that the local
DP of the rectangle alone (the oracle run on the two substrings) gives the
same score, aligned length and identity/similarity -- including low
complexity sequences, zero gap penalties (many ties) and protein.
"""
import numpy as np
import pytest

from oracle import oracle as O

BLOCK = 4  # columns per first-maximum block of the score pass


def reverse_start(s1, s2, M, o, e, bi, jb, best):
    """Reverse DP anchored on the block (no 0 floor); returns (si, sj)."""
    wend = min(jb + BLOCK - 1, len(s2))
    anc = wend - jb + 1
    r1 = s1[:bi][::-1]
    r2 = s2[:wend][::-1]
    NEG = -(1 << 40)
    n1, n2 = bi, wend
    V = [[NEG] * (n2 + 1) for _ in range(n1 + 1)]
    G = [[NEG] * (n2 + 1) for _ in range(n1 + 1)]
    H = [[NEG] * (n2 + 1) for _ in range(n1 + 1)]
    for c in range(anc):
        V[0][c] = 0
    hitr = hitc = 0
    for i in range(1, n1 + 1):
        for j in range(1, n2 + 1):
            G[i][j] = max(G[i - 1][j] - e, V[i - 1][j] - o)
            H[i][j] = max(H[i][j - 1] - e, V[i][j - 1] - o)
            V[i][j] = max(V[i - 1][j - 1] + int(M[r1[i - 1]][r2[j - 1]]), G[i][j], H[i][j])
            assert V[i][j] <= best
            if V[i][j] == best:
                hitr = max(hitr, i)
                hitc = max(hitc, j)
    if hitr == 0:
        return 1, 1, wend
    return bi - hitr + 1, wend - hitc + 1, wend


def check_pair(s1, s2, M, o, e):
    full = O.sw_literal(s1, s2, M, o, e)
    if len(s1) == 0 or len(s2) == 0:
        return
    best = int(full["score"])
    bi, bj = full["best_row"], full["best_col"]
    jb = (bj - 1) // BLOCK * BLOCK + 1
    si, sj, wend = reverse_start(s1, s2, M, o, e, bi, jb, best)
    assert 1 <= si <= bi and 1 <= sj <= bj <= wend
    sub = O.sw_literal(s1[si - 1:bi], s2[sj - 1:wend], M, o, e)
    assert int(sub["score"]) == best
    assert (sub["best_row"] + si - 1, sub["best_col"] + sj - 1) == (bi, bj)
    for k in ("columns", "identity", "similarity", "gaps"):
        assert sub[k] == full[k], (k, s1, s2, si, sj)


def mutate(rng, s, sub, indel, alphabet):
    out = []
    for ch in s:
        r = rng.random()
        if r < indel / 2:
            continue
        if r < indel:
            out.append(alphabet[rng.integers(len(alphabet))])
        out.append(alphabet[rng.integers(len(alphabet))] if rng.random() < sub else ch)
    return bytes(out)


@pytest.mark.parametrize("o,e", [(16, 4), (5, 2), (1, 1), (0, 0)])
def test_rectangle_dna(ednafull, o, e):
    rng = np.random.default_rng(11 + o)
    A = b"ACGT"
    for _ in range(150):
        n1, n2 = rng.integers(1, 45, size=2)
        s1 = bytes(A[k] for k in rng.integers(4, size=n1))
        if rng.random() < 0.5:
            s2 = mutate(rng, s1, 0.15, 0.1, A)[: max(1, n2)] or b"A"
        else:
            s2 = bytes(A[k] for k in rng.integers(4, size=n2))
        check_pair(s1, s2, ednafull, o, e)


def test_rectangle_low_complexity(ednafull):
    cases = [b"A" * 30, b"AC" * 15, b"ACGT" * 8, b"AAAACCCC" * 4, b"TTTTTTTTTA", b"GATTACA" * 4]
    for s1 in cases:
        for s2 in cases:
            for o, e in ((16, 4), (0, 0), (3, 1)):
                check_pair(s1, s2, ednafull, o, e)


def test_rectangle_protein(blosum62):
    rng = np.random.default_rng(5)
    A = b"ACDEFGHIKLMNPQRSTVWY"
    for _ in range(120):
        n1, n2 = rng.integers(1, 40, size=2)
        s1 = bytes(A[k] for k in rng.integers(20, size=n1))
        s2 = mutate(rng, s1, 0.3, 0.1, A) if rng.random() < 0.5 else bytes(A[k] for k in rng.integers(20, size=n2))
        check_pair(s1, s2 or b"W", blosum62, 10, 1)
