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


# ---- the max3 form: values biased by gap_open, three-way maxima on int16
# bit patterns read as fp16 (v_pk_maximum3_f16), the row maximum taken over
# two columns at once ----

def _i16(x):
    """Wrap to a signed 16-bit value (v_pk_sub_i16 without clamp)."""
    return ((int(x) + 0x8000) & 0xFFFF) - 0x8000


def _add_sat(a, b):
    """v_pk_add_i16 with clamp."""
    return max(-0x8000, min(0x7FFF, a + b))


def _sub_sat_u16(a, b):
    """v_pk_sub_u16 with clamp on the unsigned reading of a."""
    return max(0, (a & 0xFFFF) - b)


def _max3_f16(*xs):
    """v_pk_maximum3_f16 on one half: the operands' 16-bit patterns read as
    fp16 (numpy's float16 decodes them), IEEE 754-2019 maximum (a NaN operand
    gives a NaN); returns the result's pattern as a signed 16-bit value."""
    pats = np.array([x & 0xFFFF for x in xs], dtype=np.uint16)
    vals = pats.view(np.float16).astype(np.float64)
    if np.isnan(vals).any():
        return 0x7E00
    k = int(np.argmax(vals))
    return _i16(int(pats[k]))


def mx3_model(s1, s2, M, o, e, R, ncols_pad=0):
    """(best, bi, jb, flagged) of one pair as the max3 form computes it:
    strips of R rows (the last padded with pad rows), columns padded to a
    multiple of 4 plus `ncols_pad` (a tile's longer member), pad score 0 in
    the biased profile (a score of -gap_open)."""
    La, Lb = len(s1), len(s2)
    ncols = ((Lb + ncols_pad) + 3) & ~3
    NEG = -16384

    def sp(i, j):  # biased profile value of row i, column j (1-based)
        if i > La or j > Lb:
            return 0
        v = int(M[s1[i - 1], s2[j - 1]]) + o
        assert v >= 0
        return v

    bt = [0] * (ncols + 1)  # boundary t (= V of the row above), G biased
    bG = [0] * (ncols + 1)
    GV, GR, GJ = -1, 0, 0
    for i0 in range(0, La, R):
        tr, Hr, bm = [0] * R, [o] * R, [NEG] * R
        dtb = 0
        for j in range(1, ncols + 1):
            tu, Gu = bt[j], bG[j]
            dt = dtb
            dtb = tu
            for r in range(R):
                G = max(_i16(Gu - e), tu)
                H = _max3_f16(_sub_sat_u16(Hr[r], e), tr[r], o)
                f = _add_sat(dt, sp(i0 + r + 1, j))
                V = _max3_f16(f, H, G)
                tn = _i16(V - o)
                if j % 2 == 0:
                    bm[r] = _max3_f16(bm[r], tr[r], tn)
                dt = tr[r]
                tr[r], Hr[r] = tn, H
                tu, Gu = tn, G
            bt[j], bG[j] = tu, Gu
            if j % BLOCK == 0 or j == ncols:
                BV, BR = bm[0], 0
                for r in range(1, R):
                    if bm[r] > BV:
                        BV, BR = bm[r], r
                bm = [NEG] * R
                BR += i0 + 1
                if BV > GV or (BV == GV and BR < GR):
                    GV, GR, GJ = BV, BR, ((j - 1) & ~(BLOCK - 1)) + 1
    flagged = La == 0 or Lb == 0 or GV >= 0x7C00 - o
    return GV, GR, GJ, flagged


def mx3_cases(ednafull, blosum62):
    rng = random.Random(37)
    out = []
    for _ in range(40):
        a, b = random_related(rng, rng.randint(1, 60))
        out.append((a, b, ednafull, rng.choice([16, 10, 5, 4]), rng.choice([4, 1, 2, 0]), rng.choice([1, 4, 8])))
    aa = b"ARNDCQEGHILKMFPSTWYVX"
    for _ in range(15):
        a = bytes(rng.choice(aa) for _ in range(rng.randint(1, 50)))
        b = bytes(rng.choice(aa) for _ in range(rng.randint(1, 50)))
        out.append((a, b, blosum62, 10, 1, 8))
    out.append((b"A" * 40, b"A" * 37, ednafull, 16, 4, 8))
    out.append((b"AC" * 30, b"CA" * 31, ednafull, 4, 0, 4))  # many equal maxima, gap extension 0
    out.append((b"G", b"C", ednafull, 16, 4, 8))
    return out


@pytest.mark.parametrize("pad_cols", [0, 5, 23])
def test_max3_form_matches_the_oracle(ednafull, blosum62, pad_cols):
    """Below the threshold the max3 form gives the oracle's score, first-
    maximum row and a block holding its first-maximum column, with pad rows
    and pad columns of a longer tile member."""
    for a, b, M, o, e, R in mx3_cases(ednafull, blosum62):
        best, bi, jb, flagged = mx3_model(a, b, M, o, e, R, pad_cols)
        r = O.sw_literal(a, b, M, o, e)
        assert not flagged
        assert best == int(r["score"]), (a, b, o, e)
        assert bi == r["best_row"], (a, b, o, e, bi, r["best_row"])
        assert jb <= r["best_col"] < jb + BLOCK, (a, b, o, e, jb, r["best_col"])


def test_max3_form_flags_exactly_the_pairs_past_its_range(ednafull):
    """With scores scaled up (match 1000, mismatch -800, opening 900) the
    biased values leave the ordered fp16 range within a few dozen cells: a
    pair is flagged exactly when its true score reaches 0x7C00 - gap_open,
    and every unflagged result is exact -- also after the patterns past
    0x7BFF (infinities, NaNs) have run through later cells and strips."""
    M = np.where(np.eye(ednafull.shape[0], dtype=bool), 1000, -800).astype(np.int64)
    o, e = 900, 300
    rng = random.Random(41)
    seen = {True: 0, False: 0}
    for n in list(range(26, 36)) + [60, 90]:
        for _ in range(3):
            a = bytes(rng.choice(b"ACGT") for _ in range(n))
            b = bytearray(a)
            for _ in range(rng.randint(0, 3)):
                b[rng.randrange(n)] = rng.choice(b"ACGT")
            b = bytes(b) + bytes(rng.choice(b"ACGT") for _ in range(rng.randint(0, 9)))
            best, bi, jb, flagged = mx3_model(a, b, M, o, e, 8, 3)
            r = O.sw_literal(a, b, M, o, e)
            assert flagged == (int(r["score"]) >= 0x7C00 - o), (a, b, best, r["score"])
            seen[flagged] += 1
            if not flagged:
                assert (best, bi) == (int(r["score"]), r["best_row"])
                assert jb <= r["best_col"] < jb + BLOCK
    assert seen[True] > 5 and seen[False] > 5
