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
complexity sequences, zero gap penalties (many ties), extension above
opening, and protein.  The GPU evaluates the reverse DP only on a band (cells
with a predecessor whose value can still lie on an optimal path); the banded
restatement here must find the same rectangle as the full reverse DP.
"""
import numpy as np
import pytest

from oracle import oracle as O

# columns per first-maximum block of the score pass (16); the rectangle and
# band rules hold for any block width, 4 and 32 are checked as well
BLOCKS = (4, 16, 32)
KBLOCK = 16


def reverse_start(s1, s2, M, o, e, bi, jb, best, BLOCK):
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


PAD = -16384  # score of a pad row / column past the pair's prefix


def reverse_start_banded(s1, s2, M, o, e, bi, jb, best, R, BLOCK, live_min=None, stats=None):
    """The same reverse DP evaluated on the GPU kernel's band: strips of R
    rows, columns in 4-column groups from jlo (the first group with a live
    cell, V' >= min(0, o - e), in the row above the strip) until a group with no live
    cell past the last live column of the row above; cells outside are -inf;
    stop after a strip whose last row has no live cell or once every row
    seen so far includes one without a live cell.  Returns (si, sj)."""
    live_min = min(0, o - e) if live_min is None else live_min
    wend = min(jb + BLOCK - 1, len(s2))
    anc = wend - jb + 1
    r1 = s1[:bi][::-1]
    r2 = s2[:wend][::-1]
    NEG = -(1 << 40)
    ncols = (wend + 3) // 4 * 4
    evaluated = 0
    top_v = [NEG] * (ncols + 1)   # row above the strip: V' and the vertical-gap state
    top_g = [NEG] * (ncols + 1)
    for c in range(anc):
        top_v[c] = 0
    jlo, jtop, live_hi = 1, BLOCK, BLOCK - 1
    hitr = hitc = 0
    seen_dead = False
    for i0 in range(0, bi, R):
        rows = range(i0 + 1, i0 + R + 1)
        left_v = {i: NEG for i in rows}
        left_h = {i: NEG for i in rows}
        rmax = {i: NEG for i in rows}
        new_v = [NEG] * (ncols + 1)
        new_g = [NEG] * (ncols + 1)
        diag_top = top_v[0] if (i0 == 0 and jlo == 1) else NEG
        lo_next = hi_next = 0
        j = jlo
        while j <= ncols:
            group_live = last_live = False
            for jj in range(j, j + 4):
                up_v = top_v[jj] if jj <= jtop else NEG
                up_g = top_g[jj] if jj <= jtop else NEG
                diag = diag_top
                diag_top = up_v
                for i in rows:
                    s = int(M[r1[i - 1]][r2[jj - 1]]) if (i <= bi and jj <= wend) else PAD
                    g = max(up_g - e, up_v - o)
                    h = max(left_h[i] - e, left_v[i] - o)
                    v = max(diag + s if diag > NEG // 2 else NEG, g, h)
                    v = max(v, NEG)
                    assert v <= best
                    if v == best and i <= bi and jj <= wend:
                        hitr, hitc = max(hitr, i), max(hitc, jj)
                    rmax[i] = max(rmax[i], v)
                    group_live |= v >= live_min
                    if i == rows[-1]:
                        last_live |= v >= live_min
                    evaluated += 1
                    diag = left_v[i]
                    left_v[i], left_h[i] = v, h
                    up_v, up_g = v, g
                new_v[jj], new_g[jj] = up_v, up_g
            jend = j + 3
            if last_live:
                lo_next = lo_next or j
                hi_next = j + 3
            if not group_live and j + 3 > live_hi:
                break
            j += 4
        for i in rows:
            seen_dead |= rmax[i] < live_min or i > bi
        if lo_next == 0 or seen_dead:
            break
        jlo, live_hi, jtop = lo_next, hi_next, jend
        top_v, top_g = new_v, new_g
    if stats is not None:
        stats["evaluated"] = stats.get("evaluated", 0) + evaluated
        stats["prefix"] = stats.get("prefix", 0) + bi * wend
    if hitr == 0:
        return 1, 1
    return bi - hitr + 1, wend - hitc + 1


def check_pair(s1, s2, M, o, e):
    full = O.sw_literal(s1, s2, M, o, e)
    if len(s1) == 0 or len(s2) == 0:
        return
    best = int(full["score"])
    bi, bj = full["best_row"], full["best_col"]
    for BLOCK in BLOCKS:
        jb = (bj - 1) // BLOCK * BLOCK + 1
        si, sj, wend = reverse_start(s1, s2, M, o, e, bi, jb, best, BLOCK)
        for R in (2, 16):
            assert reverse_start_banded(s1, s2, M, o, e, bi, jb, best, R, BLOCK) == (si, sj), (R, s1, s2, o, e)
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


@pytest.mark.parametrize("o,e", [(16, 4), (5, 2), (1, 1), (0, 0), (2, 5)])
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


def _long_pairs(rng, n, lo, hi):
    A = b"ACGT"
    out = []
    for _ in range(n):
        n1, n2 = rng.integers(lo, hi, size=2)
        anc = bytes(A[k] for k in rng.integers(4, size=max(n1, n2)))
        s1 = mutate(rng, anc, 0.15, 0.03, A)[:n1]
        s2 = mutate(rng, anc, 0.15, 0.03, A)[:n2] if rng.random() < 0.3 else bytes(A[k] for k in rng.integers(4, size=n2))
        out.append((s1, s2))
    return out


def _rect(s1, s2, M, o, e, BLOCK=KBLOCK):
    full = O.sw_literal(s1, s2, M, o, e)
    best = int(full["score"])
    bi, bj = full["best_row"], full["best_col"]
    return best, bi, (bj - 1) // BLOCK * BLOCK + 1


def test_band_prunes_and_keeps_rectangle(ednafull):
    """Longer pairs (120..260 residues, as the block workloads' scale with
    related and unrelated pairs): the band evaluates a fraction of the
    prefix rectangle and still finds the unbanded DP's rectangle."""
    rng = np.random.default_rng(23)
    stats = {}
    for s1, s2 in _long_pairs(rng, 12, 120, 260):
        best, bi, jb = _rect(s1, s2, ednafull, 16, 4)
        if best == 0:
            continue
        si, sj, _ = reverse_start(s1, s2, ednafull, 16, 4, bi, jb, best, KBLOCK)
        assert reverse_start_banded(s1, s2, ednafull, 16, 4, bi, jb, best, 16, KBLOCK, stats=stats) == (si, sj)
    assert stats["evaluated"] < 0.8 * stats["prefix"], stats


def test_band_threshold_is_not_slack(ednafull):
    """The comparison catches a band that is too narrow: with liveness
    V' >= e instead of V' >= min(0, o - e) the band loses hits."""
    rng = np.random.default_rng(29)
    lost = 0
    for s1, s2 in _long_pairs(rng, 40, 30, 90):
        best, bi, jb = _rect(s1, s2, ednafull, 16, 4)
        if best == 0:
            continue
        ref = reverse_start(s1, s2, ednafull, 16, 4, bi, jb, best, KBLOCK)[:2]
        lost += reverse_start_banded(s1, s2, ednafull, 16, 4, bi, jb, best, 2, KBLOCK, live_min=4) != ref
    assert lost > 0


@pytest.mark.parametrize("o,e", [(16, 4), (0, 0), (2, 5)])
def test_prefix_rectangle_is_exact(ednafull, o, e):
    """The reverse pass's row cap gives a pair whose band is still live at
    the cap its whole prefix [1..bi] x [1..wend] as rectangle: the local DP
    of the prefix alone gives the full pair's score, first maximum and
    statistics (its values are the full DP's: a cell depends only on the
    cells above and to its left)."""
    rng = np.random.default_rng(31 + o)
    for s1, s2 in _long_pairs(rng, 30, 20, 200):
        full = O.sw_literal(s1, s2, ednafull, o, e)
        best, bi, jb = _rect(s1, s2, ednafull, o, e)
        if len(s1) == 0 or len(s2) == 0:
            continue
        wend = min(jb + KBLOCK - 1, len(s2))
        sub = O.sw_literal(s1[:bi], s2[:wend], ednafull, o, e)
        assert int(sub["score"]) == best
        assert (sub["best_row"], sub["best_col"]) == (full["best_row"], full["best_col"])
        for k in ("columns", "identity", "similarity", "gaps"):
            assert sub[k] == full[k], (k, s1, s2)


# ---- the reverse pass's max3 form: values biased by B and kept in
# [0, 0x7BFF] with unsigned saturating subtracts (a clamped value is dead and,
# the band's rows being capped, cannot climb back to life), three-way maxima
# on the int16 patterns read as fp16; the row and column maxima are taken in
# the t = V' - o domain ----

from tests.test_score_pass_model import _add_sat, _max3_f16  # noqa: E402

REV_BIAS = 16384


def _subsat_u16(a, b):
    return max(0, a - b)


def reverse_start_banded_max3(s1, s2, M, o, e, bi, jb, best, R, BLOCK, row_cap=512):
    """reverse_start_banded in the kernel's max3 arithmetic; a pair whose
    score does not fit the biased range takes its whole prefix."""
    B = REV_BIAS
    live_t = B - max(o, e)              # V' >= min(0, o - e)  <=>  t >= B - max(o, e)
    hit_t = best + B - o                 # V' == S
    wend = min(jb + BLOCK - 1, len(s2))
    if best + B > 0x7BFF:
        return 1, 1
    anc = wend - jb + 1
    r1 = s1[:bi][::-1]
    r2 = s2[:wend][::-1]
    ncols = (wend + 3) // 4 * 4
    top_t = [0] * (ncols + 1)            # -inf is 0
    top_g = [0] * (ncols + 1)
    for c in range(anc):
        top_t[c] = B - o                 # V' = 0
    jlo, jtop, live_hi = 1, BLOCK, BLOCK - 1
    hitr = hitc = 0
    seen_dead = False
    rows_done = 0
    for i0 in range(0, min(bi, row_cap), R):
        rows = range(i0 + 1, i0 + R + 1)
        left_t = {i: 0 for i in rows}
        left_h = {i: 0 for i in rows}
        rmax = {i: 0 for i in rows}
        new_t = [0] * (ncols + 1)
        new_g = [0] * (ncols + 1)
        diag_top = top_t[0] if (i0 == 0 and jlo == 1) else 0
        lo_next = hi_next = 0
        j = jlo
        while j <= ncols:
            group_live = last_live = False
            for jj in range(j, j + 4):
                up_t = top_t[jj] if jj <= jtop else 0
                up_g = top_g[jj] if jj <= jtop else 0
                diag = diag_top
                diag_top = up_t
                for i in rows:
                    sp = (int(M[r1[i - 1]][r2[jj - 1]]) + o) if (i <= bi and jj <= wend) else -0x8000
                    assert sp >= -0x8000
                    g = max(_subsat_u16(up_g, e), up_t)
                    h = max(_subsat_u16(left_h[i], e), left_t[i])
                    f = _add_sat(diag, sp)
                    v = _max3_f16(f, h, g)
                    assert 0 <= v <= 0x7BFF
                    t = _subsat_u16(v, o)
                    if t == hit_t and i <= bi and jj <= wend:
                        hitr, hitc = max(hitr, i), max(hitc, jj)
                    assert t <= hit_t
                    rmax[i] = max(rmax[i], t)
                    group_live |= t >= live_t
                    if i == rows[-1]:
                        last_live |= t >= live_t
                    diag = left_t[i]
                    left_t[i], left_h[i] = t, h
                    up_t, up_g = t, g
                new_t[jj], new_g[jj] = up_t, up_g
            jend = j + 3
            if last_live:
                lo_next = lo_next or j
                hi_next = j + 3
            if not group_live and j + 3 > live_hi:
                break
            j += 4
        rows_done = i0 + R
        for i in rows:
            seen_dead |= rmax[i] < live_t or i > bi
        if lo_next == 0 or seen_dead:
            break
        jlo, live_hi, jtop = lo_next, hi_next, jend
        top_t, top_g = new_t, new_g
    capped = bi > rows_done and not seen_dead
    if hitr == 0 or capped:
        return 1, 1
    return bi - hitr + 1, wend - hitc + 1


@pytest.mark.parametrize("R", [2, 16])
def test_max3_form_reverse_pass_finds_the_same_rectangle(ednafull, blosum62, R):
    """The biased, clamped max3 arithmetic finds the exact reverse DP's
    rectangle start on random related and unrelated pairs, low complexity
    and protein, at several penalties (and a row cap small enough to stop
    some bands, which then take the whole prefix, like the exact pass)."""
    import random
    from tests.test_oracle import random_related
    rng = random.Random(43)
    cases = []
    for _ in range(40):
        a, b = random_related(rng, rng.randint(4, 90))
        cases.append((a, b, ednafull, rng.choice([16, 10, 5, 4]), rng.choice([4, 1, 2, 0])))
    aa = b"ARNDCQEGHILKMFPSTWYV"
    for _ in range(12):
        a = bytes(rng.choice(aa) for _ in range(rng.randint(4, 70)))
        b = bytes(rng.choice(aa) for _ in range(rng.randint(4, 70)))
        cases.append((a, b, blosum62, 10, 1))
    cases += [(b"A" * 60, b"A" * 50, ednafull, 16, 4), (b"AC" * 30, b"CA" * 31, ednafull, 4, 0)]
    # multi-residue gaps in both sequences, so that the optimal paths run
    # through vertical and horizontal gap chains (extension, not opening)
    for k in range(6):
        x = bytes(rng.choice(b"ACGT") for _ in range(80))
        g1, g2 = 2 + k % 3, 3 + k % 4
        # unrelated prefixes, so that the alignment starts inside the matrix
        ja = bytes(rng.choice(b"ACGT") for _ in range(7 + k))
        jb_ = bytes(rng.choice(b"ACGT") for _ in range(12 - k))
        a = ja + x[:25] + x[25 + g1:]            # a deletion in s1
        b = jb_ + x[:50] + bytes(rng.choice(b"ACGT") for _ in range(g2)) + x[50:]   # an insertion in s2
        for o, e in ((16, 4), (10, 1), (6, 2)):
            cases += [(a, b, ednafull, o, e), (b, a, ednafull, o, e)]
    for a, b, M, o, e in cases:
        r = O.sw_literal(a, b, M, o, e)
        best, bi = int(r["score"]), r["best_row"]
        if best <= 0:
            continue
        jb = ((r["best_col"] - 1) // KBLOCK) * KBLOCK + 1
        for cap in (512, 8):
            want = reverse_start_banded(a, b, M, o, e, bi, jb, best, R, KBLOCK) if cap >= bi else None
            got = reverse_start_banded_max3(a, b, M, o, e, bi, jb, best, R, KBLOCK, row_cap=cap)
            if want is not None:
                assert got == want, (a, b, o, e, got, want)
            # whatever the cap, the rectangle holds the full DP's traceback
            si, sj = got
            wend = min(jb + KBLOCK - 1, len(b))
            sub = O.sw_literal(a[si - 1:bi], b[sj - 1:wend], M, o, e)
            assert int(sub["score"]) == best
