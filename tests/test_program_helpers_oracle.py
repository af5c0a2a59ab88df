"""Pins of the oracle's SWMS and NW ComputeDistance helpers (no GPU needed).

Salsa.SmithWatermanMS/Program.cs:417-668 and
Salsa.NeedlemanWunschTPL/Program.cs:269-300,547-700 compute their distances
from Bio.dll's aligned pair with their own helpers: PercentSimilarity counts
non-gap columns scoring > 0 only (not identical pairs that score <= 0),
Kimura2 is clamped to 10, Jukes-Cantor returns 10 where SWMPI throws, and
MinMaxNormScore = (s - minS) / (maxS - minS) with maxS the larger self score
(sum of M[c][c]) and minS 0 (SWMS) or 2*open + (|si|+|sj|)*ext (NW).  The
aligner itself is closed (parity unpinned); these tests pin the helpers on
the build's own alignment: hand-worked alignments, an independent count over
the aligned strings, and the formulas' closed values.
"""
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle import random_related


def pairs(*seqs):
    res = np.frombuffer(b"".join(seqs), dtype=np.uint8).copy()
    off = np.zeros(len(seqs) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    return res, off


def one(s1, s2, M, o, e, **kw):
    res, off = pairs(s1, s2)
    d, st = O.sw_pairs_mt(res, off, [0], [1], M, o, e, want_stats=True, **kw)
    return d[0], st[0]


def f32_fraction_i16(cnt, cols):
    r = np.float32(1.0) - np.float32(cnt) / np.float32(cols)
    return int(np.trunc(r * np.float32(32767)))


def test_positive_similarity_counts_aligned_strings(blosum62):
    rng = random.Random(5)
    aa = b"ARNDCQEGHILKMFPSTWYVXBZ*"
    for _ in range(200):
        a = bytes(rng.choice(aa) for _ in range(rng.randint(1, 60)))
        b = bytearray(a)
        for q in range(len(b)):
            if rng.random() < 0.3:
                b[q] = rng.choice(aa)
        r = O.sw_literal(a, bytes(b), blosum62, 10, 1, want_alignment=True)
        want = sum(1 for x, y in zip(r["aln1"], r["aln2"])
                   if x != "-" and y != "-" and blosum62[ord(x), ord(y)] > 0)
        assert r["pos_similarity"] == want


def test_swms_similarity_leaves_out_identical_pairs_scoring_below_one(blosum62):
    # W/W = 11, X/X = -1 under BLOSUM62: the whole pair aligns (43 > 22); SWMPI
    # counts X/X as similar (c1 == c2), SWMS does not
    d0, st0 = one(b"WWXWW", b"WWXWW", blosum62, 10, 1, nucleic=False)
    d1, st1 = one(b"WWXWW", b"WWXWW", blosum62, 10, 1, nucleic=False, program=1)
    assert st0.tolist() == [43, 5, 5, 5] and d0 == 0
    assert st1.tolist() == [43, 5, 5, 4] and d1 == f32_fraction_i16(4, 5) == 6553


def positive_mismatch_matrix():
    M = np.zeros((127, 127), dtype=np.float32)
    for a in b"ACGT":
        for b in b"ACGT":
            M[a, b] = 2 if a == b else 1
    return M


def test_swms_jukes_cantor_undefined_is_ten():
    M = positive_mismatch_matrix()
    with pytest.raises(ArithmeticError):
        one(b"AAAA", b"CCCC", M, 16, 4, dist_fn=2)
    d, _ = one(b"AAAA", b"CCCC", M, 16, 4, dist_fn=2, program=1)
    assert d == O.f64_to_i16(10.0 * 32767) == -10
    d, _ = one(b"AAAA", b"CCCC", M, 16, 4, dist_fn=2, program=1, tdist_double=True)
    assert d == 10.0


def test_swms_kimura_matches_swmpi_below_ten(ednafull):
    # the clamp at 10 cannot trigger for counts below 2^16 (the largest
    # finite value is about 0.75 ln 65535 < 8.4); both programs agree
    rng = random.Random(9)
    for _ in range(50):
        a, b = random_related(rng, rng.randint(5, 120))
        assert one(a, b, ednafull, 16, 4, dist_fn=1, tdist_double=True)[0] == \
            one(a, b, ednafull, 16, 4, dist_fn=1, tdist_double=True, program=1)[0]


def test_minmax_local(ednafull):
    # SWMS: (s - 0) * 1.0 / (maxS - 0); identical sequences: norm 1, distance 0
    d, st = one(b"ACGT", b"ACGT", ednafull, 16, 4, dist_fn=3, program=1)
    assert st[0] == 20 and d == 0
    # self scores 25 and 20, score 20: norm 0.8
    d, st = one(b"ACGTT", b"ACGT", ednafull, 16, 4, dist_fn=3, program=1)
    assert st[0] == 20 and d == int((1.0 - 20 * 1.0 / 25) * 32767) == 6553
    # the Double build has no MinMaxNormScore case
    d, _ = one(b"ACGTT", b"ACGT", ednafull, 16, 4, dist_fn=3, program=1, tdist_double=True)
    assert d == 0.0
    # an empty sequence has no local score: 0
    d, _ = one(b"", b"ACGT", ednafull, 16, 4, dist_fn=3, program=1)
    assert d == 0


def test_minmax_global(ednafull):
    res, off = pairs(b"ACGTT", b"ACGT", b"ACGT")
    d, st = O.nw_pairs_mt(res, off, [0, 1], [1, 2], ednafull, 16, 4, dist_fn=3, want_stats=True)
    for k, (i, j) in enumerate([(0, 1), (1, 2)]):
        Li, Lj = int(off[i + 1] - off[i]), int(off[j + 1] - off[j])
        self_i, self_j = 5 * Li, 5 * Lj
        s = int(st[k][0])
        min_s = 2 * 16 + (Li + Lj) * 4
        want = O.f64_to_i16((1.0 - (s - min_s) * 1.0 / (max(self_i, self_j) - min_s)) * 32767)
        assert d[k] == want
    assert int(st[1][0]) == 20 and d[1] == 0          # identical: norm 1
    assert int(st[0][0]) == 20 - 16                  # one end gap


def test_global_kimura_jukes_cantor(ednafull):
    # global alignments: ACGTACGT / ACGTACGA is 8 diagonal columns with one
    # T/A transversion; ACGTT / ACGT has one end gap and no difference
    res, off = pairs(b"ACGTACGT", b"ACGTACGA", b"ACGTT", b"ACGT")
    for fn in (1, 2):
        d, st = O.nw_pairs_mt(res, off, [0, 2], [1, 3], ednafull, 16, 4, dist_fn=fn, want_stats=True,
                              tdist_double=True)
        k2 = -0.5 * np.log(1.0 - 2.0 * 0 - 1 / 8) - 0.25 * np.log(1.0 - 2.0 * (1 / 8))
        jc = -0.75 * np.log(1.0 - (4.0 / 3.0) * (1 / 8))
        assert d[0] == (k2 if fn == 1 else jc)
        assert st[0].tolist()[1:] == ([8, 0, 1] if fn == 1 else [8, 1, 0])
        assert d[1] == 0.0 and st[1].tolist()[1:] == [4, 0, 0]
    # undefined Jukes-Cantor: the global program's artificial distance 10
    M = positive_mismatch_matrix()
    res, off = pairs(b"AAAA", b"CCCC")
    d, _ = O.nw_pairs_mt(res, off, [0], [1], M, 16, 4, dist_fn=2)
    assert d[0] == -10
