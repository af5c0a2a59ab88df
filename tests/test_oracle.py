"""Pins of the CPU oracle (oracle/swg_oracle.c) -- no GPU needed.

* brute-force enumeration (tests/bruteforce.py) on tiny inputs: scores, the
  traceback start cell, the chosen path's statistics and start position;
* the count-carrying integer DP against the literal pointer restatement on
  random and adversarial pairs;
* closed forms (self alignment, one substitution, one gap of length k);
* the distance epilogue against an independent numpy binary32 evaluation,
  exhaustively for aligned lengths <= 600 and on a random sample up to 4000.
"""
import itertools
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests import bruteforce as BF


def small_matrix(match=2, mismatch=-1, sim_pairs=()):
    M = np.zeros((127, 127), dtype=np.float32)
    for a in b"ACGT":
        for b in b"ACGT":
            M[a, b] = match if a == b else mismatch
    for a, b, v in sim_pairs:
        M[a, b] = v
        M[b, a] = v
    return M


KEYS = ("score", "best_row", "best_col", "start1", "start2", "columns", "identity", "similarity", "gaps")


def check_local(s1, s2, M, o, e):
    want = BF.local(s1, s2, M, o, e)
    got = O.sw_literal(s1, s2, M, o, e)
    for k in KEYS:
        assert got[k] == want[k], (s1, s2, o, e, k, got, want)
    carry = O.sw_carry(s1, s2, M, o, e)
    for k in ("score", "best_row", "best_col", "columns", "identity", "similarity", "gaps"):
        assert carry[k] == want[k], (s1, s2, o, e, k, carry, want)


@pytest.mark.parametrize("o,e", [(2, 1), (3, 1), (5, 2), (16, 4)])
def test_literal_matches_bruteforce_random(o, e):
    rng = random.Random(1000 + o * 10 + e)
    M = small_matrix(2, -1, sim_pairs=[(ord("A"), ord("G"), 1)])
    for _ in range(60):
        n1, n2 = rng.randint(0, 4), rng.randint(0, 4)
        s1 = bytes(rng.choice(b"ACG") for _ in range(n1))
        s2 = bytes(rng.choice(b"ACG") for _ in range(n2))
        check_local(s1, s2, M, o, e)


def test_literal_matches_bruteforce_exhaustive_binary():
    # every pair of strings over {A,C} up to length 4 (low-complexity: many ties)
    M = small_matrix(1, -1)
    words = [bytes(w) for n in range(0, 5) for w in itertools.product(b"AC", repeat=n)]
    for s1 in words[::3]:
        for s2 in words[::2]:
            check_local(s1, s2, M, 2, 1)


# Inputs on which a plausible slip in the tie rules (>= instead of > for a gap
# extension, UP/LEFT or DIAGONAL/UP order swapped) changes the reported
# alignment; each is checked against the memoised enumeration.
TIE_KATS = [
    (b"GCAAAGCCA", b"GGAACCAACC", 3, -2, 4, 1),
    (b"CAAAAACCAAC", b"ACAAACAAC", 3, -2, 4, 1),
    (b"CCACAACACCCAAACC", b"ACCACACCAAACCAA", 4, -3, 5, 1),
    (b"AAACCACA", b"AACCCAACA", 2, -1, 3, 1),
    (b"AGACGCAAG", b"CGACACGGCAACA", 3, -2, 4, 1),
    (b"CCCAAACACC", b"CCAAAACCACCAA", 3, -2, 5, 2),
    (b"ACACAC", b"CAAC", 2, -1, 3, 1),
    (b"ACACA", b"AACCAA", 2, -1, 3, 1),
    (b"ACC", b"AACCAA", 1, -1, 2, 1),
    (b"CCAAA", b"CAAAAC", 1, -1, 2, 1),
]


@pytest.mark.parametrize("s1,s2,match,mismatch,o,e", TIE_KATS)
def test_tie_kats_against_memoised_enumeration(s1, s2, match, mismatch, o, e):
    M = small_matrix(match, mismatch)
    want = BF.local_memo(s1, s2, M, o, e)
    got = O.sw_literal(s1, s2, M, o, e)
    for k in KEYS:
        assert got[k] == want[k], (k, got, want)
    car = O.sw_carry(s1, s2, M, o, e)
    for k in ("score", "columns", "identity", "gaps"):
        assert car[k] == want[k], (k, car, want)


def test_memoised_enumeration_matches_full_enumeration():
    rng = random.Random(21)
    M = small_matrix(2, -1)
    for _ in range(40):
        s1 = bytes(rng.choice(b"AC") for _ in range(rng.randint(0, 5)))
        s2 = bytes(rng.choice(b"AC") for _ in range(rng.randint(0, 5)))
        assert BF.local(s1, s2, M, 3, 1) == BF.local_memo(s1, s2, M, 3, 1)


def test_literal_matches_memoised_enumeration_random():
    rng = random.Random(23)
    for match, mismatch, o, e in [(2, -1, 3, 1), (3, -2, 4, 1), (3, -2, 5, 2), (1, -1, 2, 1)]:
        M = small_matrix(match, mismatch)
        for _ in range(40):
            al = rng.choice([b"AC", b"ACG"])
            s1 = bytes(rng.choice(al) for _ in range(rng.randint(1, 14)))
            s2 = bytes(rng.choice(al) for _ in range(rng.randint(1, 14)))
            want = BF.local_memo(s1, s2, M, o, e)
            got = O.sw_literal(s1, s2, M, o, e)
            for k in KEYS:
                assert got[k] == want[k], (s1, s2, k, got, want)


def test_tie_rules_hand_worked():
    M = small_matrix(5, -4)
    # equal maxima in one column: first row wins -> best (1,1), start (0,0)
    r = O.sw_literal(b"AA", b"A", M, 16, 4)
    assert (r["best_row"], r["best_col"], r["start1"], r["start2"]) == (1, 1, 0, 0)
    # equal maxima in one row: first column wins
    r = O.sw_literal(b"A", b"AA", M, 16, 4)
    assert (r["best_row"], r["best_col"], r["start1"], r["start2"]) == (1, 1, 0, 0)
    # all cells zero: best is (1,1) with score 0 and an empty alignment
    r = O.sw_literal(b"A", b"C", M, 16, 4)
    assert (r["score"], r["best_row"], r["best_col"], r["columns"]) == (0.0, 1, 1, 0)
    # empty sequence: Cell() stays at -Infinity, alignment empty
    r = O.sw_literal(b"", b"ACGT", M, 16, 4)
    assert r["score"] == float("-inf") and r["columns"] == 0
    # V == 0 takes precedence over DIAGONAL: "AC" vs "CA" with +1/-1, o=2,e=1:
    # V(2,1)=1 (C/C) is the first max; its diagonal predecessor V(1,0)=0 -> STOP
    M1 = small_matrix(1, -1)
    r = O.sw_literal(b"AC", b"CA", M1, 2, 1)
    assert (r["score"], r["best_row"], r["best_col"], r["columns"], r["identity"]) == (1.0, 1, 2, 1, 1)


def test_closed_forms(ednafull):
    rng = random.Random(7)
    for _ in range(20):
        L = rng.randint(20, 300)
        s = bytes(rng.choice(b"ACGT") for _ in range(L))
        r = O.sw_literal(s, s, ednafull, 16, 4)
        assert r["score"] == 5 * L and r["identity"] == L and r["columns"] == L and r["gaps"] == 0
        assert O.dist_i16(r["identity"], r["columns"]) == 0
        # one substitution in the middle: identity L-1, columns L
        k = L // 2
        t = s[:k] + bytes([b"ACGT"[(b"ACGT".index(s[k]) + 1) % 4]]) + s[k + 1:]
        r = O.sw_literal(s, t, ednafull, 16, 4)
        assert r["columns"] == L and r["identity"] == L - 1 and r["score"] == 5 * (L - 1) - 4
        # one gap of length g in s1 relative to s2: cost 16 + 4(g-1)
        g = rng.randint(1, 4)
        ins = bytes(rng.choice(b"ACGT") for _ in range(g))
        u = s[:k] + ins + s[k:]
        r = O.sw_literal(u, s, ednafull, 16, 4)
        best_no_gap = 5 * max(k, L - k)
        with_gap = 5 * L - (16 + 4 * (g - 1))
        assert r["score"] >= max(best_no_gap, with_gap)
        if with_gap > best_no_gap + 5 * g:  # the gapped alignment is the unique best
            assert r["gaps"] == g and r["columns"] == L + g and r["identity"] == L


def random_related(rng, L, sub=0.1, indel=0.03):
    s = [rng.choice(b"ACGT") for _ in range(L)]
    t = []
    for ch in s:
        u = rng.random()
        if u < indel / 2:
            continue
        if u < indel:
            t.append(rng.choice(b"ACGT"))
        t.append(rng.choice(b"ACGT") if rng.random() < sub else ch)
    return bytes(s), bytes(t)


def test_carry_matches_literal_random(ednafull, blosum62):
    rng = random.Random(11)
    cases = []
    for _ in range(300):
        if rng.random() < 0.5:
            a, b = random_related(rng, rng.randint(1, 120))
        else:
            a = bytes(rng.choice(b"ACGT") for _ in range(rng.randint(0, 120)))
            b = bytes(rng.choice(b"ACGT") for _ in range(rng.randint(0, 120)))
        cases.append((a, b, ednafull, 16, 4))
    # adversarial low-complexity sequences
    for L1, L2 in [(50, 40), (64, 64), (33, 70)]:
        for pat1, pat2 in [(b"A", b"A"), (b"AC", b"CA"), (b"ACG", b"AC"), (b"AAAC", b"AC")]:
            cases.append(((pat1 * L1)[:L1], (pat2 * L2)[:L2], ednafull, 16, 4))
            cases.append(((pat1 * L1)[:L1], (pat2 * L2)[:L2], ednafull, 5, 1))
    aa = b"ARNDCQEGHILKMFPSTWYV"
    for _ in range(150):
        a = bytes(rng.choice(aa) for _ in range(rng.randint(1, 100)))
        b = bytearray(a)
        for q in range(len(b)):
            if rng.random() < 0.3:
                b[q] = rng.choice(aa)
        cases.append((a, bytes(b), blosum62, 10, 1))
    for a, b, M, o, e in cases:
        lit = O.sw_literal(a, b, M, o, e)
        car = O.sw_carry(a, b, M, o, e)
        for k in ("score", "best_row", "best_col", "columns", "identity", "similarity", "gaps",
                  "transitions", "transversions"):
            assert lit[k] == car[k], (a, b, k, lit, car)


def numpy_pid_i16(cnt, cols):
    cnt = np.asarray(cnt, dtype=np.float32)
    cols = np.asarray(cols, dtype=np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        pct = cnt / (cols * np.float32(1.0))
        r = np.float32(1.0) - pct
        x = r * np.float32(32767)
    out = np.where(np.isnan(x), 0, np.trunc(np.nan_to_num(x))).astype(np.int64)
    return out.astype(np.int16)


def test_epilogue_int16_matches_numpy_binary32():
    cols, cnt = [], []
    for c in range(0, 601):
        for k in range(0, c + 1):
            cols.append(c); cnt.append(k)
    rng = np.random.default_rng(3)
    cs = rng.integers(601, 4001, size=20000)
    ks = (rng.random(20000) * (cs + 1)).astype(np.int64)
    cols += cs.tolist(); cnt += ks.tolist()
    want = numpy_pid_i16(cnt, cols)
    got = np.array([O.dist_i16(k, c) for k, c in zip(cnt, cols)], dtype=np.int16)
    assert np.array_equal(got, want)
    assert O.dist_i16(3, 7) == 18723  # x64 binary32 chain (an x87 chain gives 18724)
    assert O.dist_i16(0, 0) == 0      # NaN -> cvttss2si 0x80000000 -> low 16 bits 0


def test_epilogue_double():
    assert np.isnan(O.dist_f64(0, 0))
    assert O.dist_f64(3, 7) == float(np.float32(1) - np.float32(3) / np.float32(7))
    assert O.dist_f64(5, 5) == 0.0


def test_kimura_and_conversions():
    assert O.kimura2(10, 0, 0, 0) == 0.0
    # 1 - (2P + Q) <= 0 -> artificial distance 10
    assert O.kimura2(10, 0, 5, 0) == 10.0
    assert O.f64_to_i16(10.0 * 32767) == -10  # 327670 wraps in the low 16 bits
    P, Q = 2 / 10, 1 / 10
    assert abs(O.kimura2(10, 0, 2, 1) - (-0.5 * np.log(1 - 2 * P - Q) - 0.25 * np.log(1 - 2 * Q))) < 1e-15


def test_global_matches_bruteforce():
    rng = random.Random(5)
    M = small_matrix(2, -1, sim_pairs=[(ord("A"), ord("G"), 1)])
    for _ in range(80):
        n1, n2 = rng.randint(0, 4), rng.randint(0, 4)
        s1 = bytes(rng.choice(b"ACG") for _ in range(n1))
        s2 = bytes(rng.choice(b"ACG") for _ in range(n2))
        for o, e in [(3, 1), (2, 1)]:
            want = BF.global_(s1, s2, M, o, e)
            got = O.nw_literal(s1, s2, M, o, e)
            for k in ("score", "columns", "identity", "similarity", "gaps", "nw_similarity"):
                assert got[k] == want[k], (s1, s2, o, e, k, got, want)
            if want["nw_region_len"] > 0:
                assert got["nw_region_len"] == want["nw_region_len"]
                assert got["nw_region_identity"] == want["nw_region_identity"]


def test_symbol_range_is_enforced(ednafull):
    with pytest.raises(IndexError):
        O.sw_literal(bytes([200]), b"A", ednafull, 16, 4)
