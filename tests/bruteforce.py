"""Brute-force alignment enumerators used to pin the CPU oracle.

A different formulation from the reference's DP: every alignment path is
enumerated explicitly, scored with affine gap runs (a run of k gap columns
costs open + (k-1)*ext), and the reported alignment is chosen by a global
rule instead of pointers:

* local: the end cell is the first cell, in row-major order, whose best
  score max(0, best path ending there) is maximal; among the optimal paths
  ending there the one whose backward step sequence is lexicographically
  smallest under STOP < DIAGONAL < UP(k) < LEFT(k), shorter gap runs first.
* global: the path from (m, n) to (0, 0), same selection rule; border runs
  along row 0 / column 0 are single gap runs.

These rules are the closed-form reading of the reference's tie order
(SmithWatermanGotoh.cs:162,177 strict '>' extension, :193-208 pointer order,
:211 first strict maximum); the tests compare the oracle against them on
inputs small enough to enumerate.
"""

STOP, DIAG, UP, LEFT = 0, 1, 2, 3


def _diag_stats(c1, c2, M):
    ident = 1 if c1 == c2 else 0
    sim = 1 if (c1 == c2 or M[c1][c2] > 0) else 0
    return ident, sim


def _paths_local(s1, s2, M, o, e, i, j, last):
    """Yield (score, key, start, columns, identity, similarity, gaps) for every
    backward path from cell (i, j)."""
    yield (0, ((STOP, 0),), (i, j), 0, 0, 0, 0)
    if i >= 1 and j >= 1:
        c1, c2 = s1[i - 1], s2[j - 1]
        s = M[c1][c2]
        di, ds = _diag_stats(c1, c2, M)
        for sc, key, st, cols, idn, sim, gp in _paths_local(s1, s2, M, o, e, i - 1, j - 1, DIAG):
            yield (sc + s, ((DIAG, 0),) + key, st, cols + 1, idn + di, sim + ds, gp)
    if j >= 1 and last != UP:
        for k in range(1, i + 1):
            cost = o + (k - 1) * e
            for sc, key, st, cols, idn, sim, gp in _paths_local(s1, s2, M, o, e, i - k, j, UP):
                yield (sc - cost, ((UP, k),) + key, st, cols + k, idn, sim, gp + k)
    if i >= 1 and last != LEFT:
        for k in range(1, j + 1):
            cost = o + (k - 1) * e
            for sc, key, st, cols, idn, sim, gp in _paths_local(s1, s2, M, o, e, i, j - k, LEFT):
                yield (sc - cost, ((LEFT, k),) + key, st, cols + k, idn, sim, gp + k)


def _best(paths):
    best = None
    for p in paths:
        if best is None or p[0] > best[0] or (p[0] == best[0] and p[1] < best[1]):
            best = p
    return best


def local(s1, s2, M, o, e):
    """Reference semantics of SmithWatermanGotoh.Align by enumeration.
    s1/s2: bytes; M: 127x127 table indexable by byte codes."""
    n1, n2 = len(s1), len(s2)
    if n1 == 0 or n2 == 0:
        return dict(score=float("-inf"), best_row=0, best_col=0, start1=0, start2=0, columns=0,
                    identity=0, similarity=0, gaps=0)
    bestv, bestcell, bestpath = None, None, None
    for i in range(1, n1 + 1):
        for j in range(1, n2 + 1):
            p = _best(_paths_local(s1, s2, M, o, e, i, j, None))
            if bestv is None or p[0] > bestv:
                bestv, bestcell, bestpath = p[0], (i, j), p
    sc, key, st, cols, idn, sim, gp = bestpath
    return dict(score=float(sc), best_row=bestcell[0], best_col=bestcell[1], start1=st[0], start2=st[1],
                columns=cols, identity=idn, similarity=sim, gaps=gp)


def local_memo(s1, s2, M, o, e):
    """Same selection rule as local(), evaluated with memoisation over
    (cell, previous step type) so that inputs of ~20 residues are feasible.
    (Optimal substructure: the best path from a cell is its best first step
    followed by the best path from the landing cell under the no-repeat
    restriction; lexicographic order is applied to whole step sequences.)"""
    n1, n2 = len(s1), len(s2)
    if n1 == 0 or n2 == 0:
        return dict(score=float("-inf"), best_row=0, best_col=0, start1=0, start2=0, columns=0,
                    identity=0, similarity=0, gaps=0)
    memo = {}

    def best(i, j, last):
        k_ = (i, j, last)
        if k_ in memo:
            return memo[k_]
        cands = [(0, ((STOP, 0),), (i, j), 0, 0, 0, 0)]
        if i >= 1 and j >= 1:
            c1, c2 = s1[i - 1], s2[j - 1]
            di, ds = _diag_stats(c1, c2, M)
            sc, key, st, cols, idn, sim, gp = best(i - 1, j - 1, DIAG)
            cands.append((sc + M[c1][c2], ((DIAG, 0),) + key, st, cols + 1, idn + di, sim + ds, gp))
        if j >= 1 and last != UP:
            for k in range(1, i + 1):
                sc, key, st, cols, idn, sim, gp = best(i - k, j, UP)
                cands.append((sc - o - (k - 1) * e, ((UP, k),) + key, st, cols + k, idn, sim, gp + k))
        if i >= 1 and last != LEFT:
            for k in range(1, j + 1):
                sc, key, st, cols, idn, sim, gp = best(i, j - k, LEFT)
                cands.append((sc - o - (k - 1) * e, ((LEFT, k),) + key, st, cols + k, idn, sim, gp + k))
        r = _best(cands)
        memo[k_] = r
        return r

    bestv, bestcell, bestpath = None, None, None
    for i in range(1, n1 + 1):
        for j in range(1, n2 + 1):
            p = best(i, j, None)
            if bestv is None or p[0] > bestv:
                bestv, bestcell, bestpath = p[0], (i, j), p
    sc, key, st, cols, idn, sim, gp = bestpath
    return dict(score=float(sc), best_row=bestcell[0], best_col=bestcell[1], start1=st[0], start2=st[1],
                columns=cols, identity=idn, similarity=sim, gaps=gp)


def _paths_global(s1, s2, M, o, e, i, j, last):
    if i == 0 and j == 0:
        yield (0, ((STOP, 0),), 0, 0, 0, 0, ())
        return
    if j == 0:  # border: one vertical run to the origin
        if last == UP:
            return
        yield (-(o + (i - 1) * e), ((UP, i), (STOP, 0)), i, 0, 0, i, (("U", i),))
        return
    if i == 0:
        if last == LEFT:
            return
        yield (-(o + (j - 1) * e), ((LEFT, j), (STOP, 0)), j, 0, 0, j, (("L", j),))
        return
    c1, c2 = s1[i - 1], s2[j - 1]
    s = M[c1][c2]
    di, ds = _diag_stats(c1, c2, M)
    for sc, key, cols, idn, sim, gp, runs in _paths_global(s1, s2, M, o, e, i - 1, j - 1, DIAG):
        yield (sc + s, ((DIAG, 0),) + key, cols + 1, idn + di, sim + ds, gp, (("D", (c1, c2)),) + runs)
    if last != UP:
        for k in range(1, i + 1):
            cost = o + (k - 1) * e
            for sc, key, cols, idn, sim, gp, runs in _paths_global(s1, s2, M, o, e, i - k, j, UP):
                yield (sc - cost, ((UP, k),) + key, cols + k, idn, sim, gp + k, (("U", k),) + runs)
    if last != LEFT:
        for k in range(1, j + 1):
            cost = o + (k - 1) * e
            for sc, key, cols, idn, sim, gp, runs in _paths_global(s1, s2, M, o, e, i, j - k, LEFT):
                yield (sc - cost, ((LEFT, k),) + key, cols + k, idn, sim, gp + k, (("L", k),) + runs)


def global_(s1, s2, M, o, e):
    """Global variant by enumeration; also returns the aligned-region identity
    of Salsa.NeedlemanWunschTPL/Program.cs:547-571 computed from the runs."""
    n1, n2 = len(s1), len(s2)
    best = None
    for p in _paths_global(s1, s2, M, o, e, n1, n2, None):
        if best is None or p[0] > best[0] or (p[0] == best[0] and p[1] < best[1]):
            best = p
    sc, key, cols, idn, sim, gp, runs = best
    # forward column list: runs are stored from the end backwards
    colsA, colsB = [], []
    ii, jj = 0, 0
    fwd = []
    for kind, val in reversed(runs):
        fwd.append((kind, val))
    for kind, val in fwd:
        if kind == "D":
            colsA.append(val[0]); colsB.append(val[1])
        elif kind == "U":
            for _ in range(val):
                colsA.append(s1[ii]); colsB.append(None); ii += 1
            continue
        else:
            for _ in range(val):
                colsA.append(None); colsB.append(s2[jj]); jj += 1
            continue
        ii += 1; jj += 1
    fa = next((q for q, c in enumerate(colsA) if c is not None), -1)
    fb = next((q for q, c in enumerate(colsB) if c is not None), -1)
    la = max((q for q, c in enumerate(colsA) if c is not None), default=-1)
    lb = max((q for q, c in enumerate(colsB) if c is not None), default=-1)
    first, last = max(fa, fb), min(la, lb)
    rid = sum(1 for q in range(first, last + 1) if q >= 0 and colsA[q] is not None and colsA[q] == colsB[q])
    nwsim = sum(1 for a, b in zip(colsA, colsB) if a is not None and b is not None and M[a][b] > 0)
    return dict(score=float(sc), columns=cols, identity=idn, similarity=sim, gaps=gp,
                nw_region_len=last - first + 1, nw_region_identity=rid, nw_similarity=nwsim)
