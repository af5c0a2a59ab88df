"""Pins of the oracle's Kimura2 and Jukes-Cantor distances (no GPU needed).

PairwiseAlignment.ComputeKimuraDistance / ComputeJukesCantorDistance
(Salsa.CoreTPL/Bio/Algorithms/Alignment.cs:254-348) count, over the aligned
columns, transitions (A<->G, C<->T), transversions (other differing pairs)
and gaps (differing pairs where one character is '-').  The counts are pinned
by hand-worked alignments whose optimal local path is forced, by the two
independent oracle implementations (pointer traceback vs carried counts) on
random inputs that include '-' residues, and the formulas by their closed
values (log 1 = 0, the artificial distance 10, the undefined Jukes-Cantor
case) and by the x64 conversions to Int16.
"""
import math
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle import random_related


def seqset(*seqs):
    res = np.frombuffer(b"".join(seqs), dtype=np.uint8).copy()
    off = np.zeros(len(seqs) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    return res, off


def dist(s1, s2, M, o, e, dist_fn, double=True):
    res, off = seqset(s1, s2)
    d, st = O.sw_pairs_mt(res, off, [0], [1], M, o, e, nucleic=True, dist_fn=dist_fn, tdist_double=double,
                          want_stats=True)
    return d[0], st[0]


BASE = b"ACGTACGTAC" + b"GTACGTACGT"


# (s1, s2, gap_open, gap_ext, columns, kimura gaps, transitions, transversions);
# each alignment is forced: the full-length path (with the gap where marked)
# scores above every alternative under EDNAFULL
KATS = [
    (b"AAAAGAAAA", b"AAAAAAAAA", 16, 4, 9, 0, 1, 0),            # one G/A transition
    (b"AAAACAAAA", b"AAAAAAAAA", 16, 4, 9, 0, 0, 1),            # one C/A transversion
    (b"AACAAGAATAA", b"AAAAAAAAAAA", 16, 4, 11, 0, 1, 2),       # C/A, G/A, T/A
    # one inserted T (a gap column) and one G->A transition: 19 matches - 4 - 5
    (BASE, BASE[:10] + b"T" + BASE[10:14] + b"A" + BASE[15:], 5, 1, 21, 1, 1, 0),
    # '-' as a residue: the ('-', G) diagonal column is a Kimura gap, not a difference
    (b"AC-TACGTAC", b"ACGTACGTAC", 16, 4, 10, 1, 0, 0),
]


@pytest.mark.parametrize("s1,s2,o,e,cols,kg,ts,tv", KATS)
def test_counts_of_forced_alignments(ednafull, s1, s2, o, e, cols, kg, ts, tv):
    for r in (O.sw_literal(s1, s2, ednafull, o, e), O.sw_carry(s1, s2, ednafull, o, e)):
        assert (r["columns"], r["kimura_gaps"], r["transitions"], r["transversions"]) == (cols, kg, ts, tv)
    # the gap column of the literal traceback is also a Kimura gap; a '-' residue is not a traceback gap
    lit = O.sw_literal(s1, s2, ednafull, o, e)
    assert lit["gaps"] == (1 if b"-" not in s1 and kg else 0)


@pytest.mark.parametrize("s1,s2,o,e,cols,kg,ts,tv", KATS)
def test_distances_of_forced_alignments(ednafull, s1, s2, o, e, cols, kg, ts, tv):
    den = cols - kg
    P, Q = ts / den, tv / den
    k2 = -0.5 * math.log(1.0 - 2.0 * P - Q) - 0.25 * math.log(1.0 - 2.0 * Q)
    jc = -0.75 * math.log(1.0 - (4.0 / 3.0) * ((ts + tv) / den))
    d, st = dist(s1, s2, ednafull, o, e, 1)
    assert d == k2 and st.tolist()[1:] == [den, ts, tv]
    d, st = dist(s1, s2, ednafull, o, e, 2)
    assert d == jc and st.tolist()[1:] == [den, ts + tv, 0]
    d16, _ = dist(s1, s2, ednafull, o, e, 1, double=False)
    assert d16 == int(k2 * 32767)
    d16, _ = dist(s1, s2, ednafull, o, e, 2, double=False)
    assert d16 == int(jc * 32767)


def test_closed_values():
    # no differences: -0.5 log 1 - 0.25 log 1 is -0.0 (the sign of zero the CLR produces too)
    assert math.copysign(1.0, O.kimura2(10, 0, 0, 0)) == -1.0 and O.kimura2(10, 0, 0, 0) == 0.0
    assert O.jukes_cantor(10, 3, 0) == 0.0
    # 1 - (2P + Q) <= 0 or 1 - 2Q <= 0: the artificial distance 10 (Alignment.cs:289-299)
    assert O.kimura2(10, 0, 5, 0) == 10.0
    assert O.kimura2(10, 0, 3, 4) == 10.0
    assert O.kimura2(10, 0, 0, 5) == 10.0
    assert O.kimura2(10, 0, 2, 5) == 10.0
    assert O.f64_to_i16(10.0 * 32767) == -10  # 327670 wraps in the low 16 bits
    # gaps leave the denominator: 1 transition over 10 - 2 columns
    assert O.kimura2(10, 2, 1, 0) == -0.5 * math.log(1.0 - 2.0 * (1 / 8)) - 0.25 * math.log(1.0)
    # d = 1 - 4/3 * 3/4 = 0: the reference throws (Alignment.cs:340-345)
    with pytest.raises(ArithmeticError):
        O.jukes_cantor(8, 0, 6)
    with pytest.raises(ArithmeticError):
        O.jukes_cantor(8, 0, 7)
    # empty alignment: 0/0 -> NaN, no exception, Int16 0
    assert math.isnan(O.kimura2(0, 0, 0, 0)) and math.isnan(O.jukes_cantor(0, 0, 0))
    assert O.f64_to_i16(float("nan")) == 0


def test_jukes_cantor_undefined_in_block_driver():
    # mismatches scoring above zero make an all-different local alignment
    M = np.zeros((127, 127), dtype=np.float32)
    for a in b"ACGT":
        for b in b"ACGT":
            M[a, b] = 1
    with pytest.raises(ArithmeticError):
        dist(b"AAAA", b"CCCC", M, 16, 4, 2)
    d, st = dist(b"AAAA", b"GGGG", M, 16, 4, 1)  # all transitions: P = 1
    assert d == 10.0 and st.tolist()[1:] == [4, 4, 0]


def test_carry_matches_literal_with_gap_residues(ednafull):
    rng = random.Random(23)
    for _ in range(300):
        a, b = random_related(rng, rng.randint(1, 90))
        a, b = bytearray(a), bytearray(b)
        for s in (a, b):
            for q in range(len(s)):
                if rng.random() < 0.05:
                    s[q] = ord("-")
        lit = O.sw_literal(bytes(a), bytes(b), ednafull, 16, 4)
        car = O.sw_carry(bytes(a), bytes(b), ednafull, 16, 4)
        for k in ("score", "columns", "gaps", "transitions", "transversions", "kimura_gaps", "identity"):
            assert lit[k] == car[k], (bytes(a), bytes(b), k, lit, car)
