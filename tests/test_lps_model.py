"""CPU model of the lanes-per-pair int16 score pass (lps16_kernel, 16 or 32
lanes per pair),
checked against the oracle's literal aligner (SmithWatermanGotoh.cs:119-219).

The model follows the kernel's data flow step by step: 16 or 32 lanes of R rows per
stripe, lane k computing column t - k + 1 at step t with the row above and
the column's codes taken from lane k-1's previous step (lane 0 from the
boundary row of the previous stripe), pre-columns and pad columns / rows
scored through the byte table whose pad entries are 0 (a score of
-gap_open), the first maximum folded per lane every 16 steps into (value,
row, block of 16 columns starting at the lane's first column of the window,
clamped to 1) and the lanes merged at the end (larger value, then earlier
row).  It must give the oracle's score, its first-maximum row, and a block
holding its first-maximum column -- the contract of the score pass that the
reverse pass and the re-run rely on.  The max3 form (values biased by
gap_open, three-way maxima on the int16 patterns read as fp16, the row
maximum over two steps at once) is modelled too."""
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle import random_related
from tests.test_score_pass_model import _add_sat, _i16, _max3_f16, _sub_sat_u16

BLOCK = 16  # kFirstMaxBlock: columns per first-maximum block, steps per fold
PAD16 = -16384


def sat16(x):
    return max(-32768, min(32767, x))


def lps_model(s1, s2, M, o, e, R, extra=b"", LANES=16, mx3=False):
    """(best, row, jb) of pair (s1, s2) as lps16_kernel<R> computes it, for
    one half (the other half of a lane runs the same recurrence on its own
    sequence).  Codes as the library assigns them: the column symbols
    first, then the others (and `extra` symbols of the matrix, as a server
    codes them), then the pad code; the column selector is min(code, 12)."""
    cols = sorted(set(s2))
    rest = sorted((set(s1) | set(extra)) - set(cols))
    codes = cols + rest
    code_of = {c: k for k, c in enumerate(codes)}
    pad = len(codes)
    assert len(cols) <= 7 and (pad < 8 or pad >= 12)

    def tab(c1, c2):
        c2 = min(c2, 12)
        if c1 == pad or c2 >= pad or c2 >= 8:
            return 0
        v = int(M[codes[c1], codes[c2]]) + o
        assert 0 <= v <= 255
        return v

    La, Lb = len(s1), len(s2)
    q1 = [code_of[c] for c in s1]
    q2 = [code_of[c] for c in s2]
    nsteps = (Lb + (LANES - 1) + 3) & ~3
    NEG = PAD16
    # t = V - o, G and H' of DP row 0 / column 0 (max3 form: biased by o)
    T0, G0, H0 = (0, 0, o) if mx3 else (-o, NEG, 0)
    # boundary row (t, G) per column 1 .. nsteps + 8: DP row 0
    bnd_t = [T0] * (nsteps + 9)
    bnd_G = [G0] * (nsteps + 9)
    GV = [-1] * LANES
    GR = [0] * LANES
    GJ = [0] * LANES
    for i0 in range(0, La, LANES * R):
        rowcode = [[q1[i0 + k * R + r] if i0 + k * R + r < La else pad for r in range(R)] for k in range(LANES)]
        tr = [[T0] * R for _ in range(LANES)]
        Hr = [[H0] * R for _ in range(LANES)]
        bm = [[NEG] * R for _ in range(LANES)]
        ot = [T0] * LANES
        oG = [G0] * LANES
        pt = [T0] * LANES
        sel = [pad] * LANES
        new_bnd_t, new_bnd_G = list(bnd_t), list(bnd_G)
        for tt in range(nsteps):
            # DPP row shift: every lane reads lane k-1's values of the previous step
            prev_sel, prev_ot, prev_oG = list(sel), list(ot), list(oG)
            for k in range(LANES):
                if k == 0:
                    j = tt + 1
                    sel[k] = q2[j - 1] if j <= Lb else pad
                    tu, Gu = bnd_t[tt + 1], bnd_G[tt + 1]
                else:
                    sel[k] = prev_sel[k - 1]
                    tu, Gu = prev_ot[k - 1], prev_oG[k - 1]
                dt = pt[k]
                pt[k] = tu
                for r in range(R):
                    sp = tab(rowcode[k][r], sel[k])
                    if mx3:
                        G = max(_i16(Gu - e), tu)
                        H = _max3_f16(_sub_sat_u16(Hr[k][r], e), tr[k][r], o)
                        f = _add_sat(dt, sp)
                        V = _max3_f16(f, H, G)
                        tn = _i16(V - o)
                        if tt % 2 == 1:
                            bm[k][r] = _max3_f16(bm[k][r], tr[k][r], tn)
                    else:
                        G = max(Gu - e, tu)
                        H = max(max(Hr[k][r] - e, 0), tr[k][r])
                        f = sat16(dt + sp)
                        V = max(f, H, G)
                        tn = V - o
                        bm[k][r] = max(bm[k][r], V)
                    dt = tr[k][r]
                    tr[k][r] = tn
                    Hr[k][r] = H
                    tu, Gu = tn, G
                ot[k], oG[k] = tu, Gu
                if k == LANES - 1 and tt >= LANES - 1:
                    new_bnd_t[tt - (LANES - 2)] = ot[k]
                    new_bnd_G[tt - (LANES - 2)] = oG[k]
            if (tt & (BLOCK - 1)) == BLOCK - 1 or tt + 1 >= nsteps:
                start = tt & ~(BLOCK - 1)
                for k in range(LANES):
                    jb = max(1, start - k + 1)
                    BV, BR = bm[k][0], 0
                    for r in range(1, R):
                        if bm[k][r] > BV:
                            BV, BR = bm[k][r], r
                    BR += i0 + k * R + 1
                    if BV > GV[k] or (BV == GV[k] and BR < GR[k]):
                        GV[k], GR[k], GJ[k] = BV, BR, jb
                    bm[k] = [NEG] * R
        bnd_t, bnd_G = new_bnd_t, new_bnd_G
    best, brow, bcol = -1, 0, 0
    for k in range(LANES):
        if GV[k] > best or (GV[k] == best and GR[k] < brow):
            best, brow, bcol = GV[k], GR[k], GJ[k]
    return best, brow, bcol


def cases(ednafull):
    rng = random.Random(7)
    out = []
    for _ in range(50):
        a, b = random_related(rng, rng.randint(1, 90))
        out.append((a, b, rng.choice([16, 10, 5]), rng.choice([4, 1, 2])))
    for _ in range(10):
        a = bytes(rng.choice(b"ACGT") for _ in range(rng.randint(1, 80)))
        b = bytes(rng.choice(b"ACGT") for _ in range(rng.randint(1, 80)))
        out.append((a, b, 16, 4))
    out.append((b"A" * 70, b"A" * 37, 16, 4))     # saturating-free long diagonal, ties along rows
    out.append((b"ACGT" * 18, b"TGCA" * 9, 4, 3))   # the smallest eligible opening (mismatch -4)
    out.append((b"AC" * 40, b"CA" * 40, 5, 1))    # many equal maxima
    out.append((b"G", b"C", 16, 4))                # every cell 0: the first maximum is (1, 1)
    return out


@pytest.mark.parametrize("mx3", [False, True])
@pytest.mark.parametrize("R,extra,lanes", [(1, b"", 16), (2, b"", 16), (8, b"", 16), (2, b"BDHKMRSUVWY", 16),
                                           (1, b"", 32), (8, b"", 32)])
def test_lps_model_matches_the_oracle(ednafull, R, extra, lanes, mx3):
    for a, b, o, e in cases(ednafull):
        best, bi, jb = lps_model(a, b, ednafull, o, e, R, extra, LANES=lanes, mx3=mx3)
        r = O.sw_literal(a, b, ednafull, o, e)
        assert best == int(r["score"]), (a, b, o, e)
        assert bi == r["best_row"], (a, b, o, e, bi, r["best_row"])
        assert 1 <= jb <= r["best_col"] < jb + BLOCK, (a, b, o, e, jb, r["best_col"])
