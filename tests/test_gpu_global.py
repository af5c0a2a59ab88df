"""GPU parity of the global (Needleman-Wunsch) variant against the oracle.

The distance is the one of Salsa.NeedlemanWunschTPL/Program.cs:269-300:
nucleic = identity over the aligned region [first non-gap, last non-gap]
(:547-571), protein = positive-score non-gap columns over all columns
(:583-604).  The alignment itself follows oracle_nw_literal (the reference's
global aligner ships only as a compiled library, so its tie order is this
oracle's own definition, pinned against brute-force enumeration).
"""
import sys

import numpy as np
import pytest

import csharp_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = O.cpu_threads()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_nw_block(seqs, r0, r1, c0, c1, diag_lower, M, o, e, nucleic=True, tdouble=False):
    rows, cols = r1 - r0 + 1, c1 - c0 + 1
    a, b = [], []
    for i in range(r0, r1 + 1):
        for j in range(c0, c1 + 1):
            if diag_lower and j >= i:
                continue
            a.append(i); b.append(j)
    d, st = O.nw_pairs_mt(seqs.residues, seqs.offsets, a, b, M, o, e, nucleic=nucleic, tdist_double=tdouble,
                          threads=THREADS, want_stats=True)
    out = np.zeros((rows, cols), dtype=np.float64 if tdouble else np.int16)
    stats = np.zeros((rows, cols, 4), dtype=np.int32)
    a = np.asarray(a) - r0
    b = np.asarray(b) - c0
    out[a, b] = d
    stats[a, b] = st
    if diag_lower:
        out[b, a] = d
        stats[b, a] = stats[a, b]
    return out, stats


def assert_same(got, want, gst, wst, what):
    if got.dtype == np.float64:
        same = (got == want) | (np.isnan(got) & np.isnan(want))
    else:
        same = got == want
    if not same.all():
        idx = np.argwhere(~same)[:5]
        raise AssertionError("%s: %d mismatches, first %s got %s want %s stats got %s want %s" % (
            what, (~same).sum(), idx.tolist(), [got[tuple(i)] for i in idx], [want[tuple(i)] for i in idx],
            [gst[tuple(i)].tolist() for i in idx], [wst[tuple(i)].tolist() for i in idx]))
    bad = np.argwhere(~(gst == wst).all(axis=-1))
    assert bad.size == 0, "%s: stats differ at %s got %s want %s" % (
        what, bad[:3].tolist(), [gst[tuple(i)].tolist() for i in bad[:3]], [wst[tuple(i)].tolist() for i in bad[:3]])


@pytest.fixture(scope="module")
def dna_set():
    s = ca.synth_dna(21, 150, 10, 0.12, 0.03, 0, [1, 240])
    extra = [b"", b"A", b"ACGT" * 30, b"A" * 90, b"T" * 200, b"GATTACA", b"C" * 5 + b"ACGT" * 20 + b"G" * 40]
    return ca.SequenceSet.from_strings([s.seq(i) for i in range(len(s))] + extra)


@pytest.fixture(scope="module")
def protein_set():
    s = ca.synth_protein(22, 100, 2, [100, 0.5, 1, 300])
    extra = [b"", b"W", b"XXBZ*AR", b"ACDEFGHIKLMNPQRSTVWY" * 4]
    return ca.SequenceSet.from_strings([s.seq(i) for i in range(len(s))] + extra)


@pytest.mark.parametrize("schedule,strip", [(ca.SCHED_PERSISTENT, 0), (ca.SCHED_PLAIN, 0),
                                            (ca.SCHED_PERSISTENT_I32, 8), (ca.SCHED_PERSISTENT_I32, 16)])
def test_nw_offdiag_nucleic(dna_set, ednafull, schedule, strip):
    say("nw offdiag nucleic schedule=%d strip=%d" % (schedule, strip))
    al = ca.Aligner(ednafull, 16, 4, mode=ca.MODE_NW_GLOBAL, schedule=schedule)
    al.set_schedule(schedule, strip)
    al.set_sequences(dna_set)
    n = len(dna_set)
    r0, r1, c0, c1 = 0, 69, 70, n - 1
    got, gst, pairs = al.compute_block(r0, r1, c0, c1, False, stats=True)
    want, wst = oracle_nw_block(dna_set, r0, r1, c0, c1, False, ednafull, 16, 4)
    assert pairs == (r1 - r0 + 1) * (c1 - c0 + 1)
    assert_same(got, want, gst, wst, "nw offdiag")
    al.close()


def test_nw_diagonal_lower_mirror(dna_set, ednafull):
    say("nw diag lower-mirror")
    al = ca.Aligner(ednafull, 16, 4, mode=ca.MODE_NW_GLOBAL)
    al.set_sequences(dna_set)
    n = len(dna_set)
    r0, r1 = 90, n - 1
    got, gst, _ = al.compute_block(r0, r1, r0, r1, True, stats=True)
    want, wst = oracle_nw_block(dna_set, r0, r1, r0, r1, True, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "nw diag")
    al.close()


@pytest.mark.parametrize("tdouble", [False, True])
def test_nw_protein(protein_set, blosum62, tdouble):
    say("nw protein tdouble=%d" % tdouble)
    al = ca.Aligner(blosum62, 11, 1, alignment_type=ca.ALIGN_PROTEIN, mode=ca.MODE_NW_GLOBAL,
                    tdistance=ca.TDIST_DOUBLE if tdouble else ca.TDIST_INT16)
    al.set_sequences(protein_set)
    n = len(protein_set)
    got, gst, _ = al.compute_block(0, 49, 50, n - 1, False, stats=True)
    want, wst = oracle_nw_block(protein_set, 0, 49, 50, n - 1, False, blosum62, 11, 1, nucleic=False,
                                tdouble=tdouble)
    assert_same(got, want, gst, wst, "nw protein")
    al.close()


def test_nw_long_pairs(ednafull):
    say("nw long pairs")
    s = ca.synth_dna(23, 6, 2, 0.15, 0.04, 0, [2500, 4100])
    al = ca.Aligner(ednafull, 16, 4, mode=ca.MODE_NW_GLOBAL)
    al.set_sequences(s)
    got, gst, _ = al.compute_block(0, 5, 0, 5, True, stats=True)
    want, wst = oracle_nw_block(s, 0, 5, 0, 5, True, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "nw long")
    al.close()


def test_nw_rejects_score_pass_schedules(dna_set, ednafull):
    al = ca.Aligner(ednafull, 16, 4, mode=ca.MODE_NW_GLOBAL)
    al.set_sequences(dna_set)
    for sch in (ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT_LPP16):
        al.set_schedule(sch, 0)
        with pytest.raises(ca.SwgError):
            al.compute_block(0, 3, 4, 7, False)
    al.close()


@pytest.mark.parametrize("kind,dist,strip", [("protein", ca.DIST_PERCENT_IDENTITY, 16),
                                             ("protein", ca.DIST_PERCENT_IDENTITY, 8),
                                             ("dna", ca.DIST_PERCENT_IDENTITY, 16),
                                             ("dna", ca.DIST_JUKES_CANTOR, 16),
                                             ("dna", ca.DIST_KIMURA2, 8)])
def test_nw_without_statistics(dna_set, protein_set, ednafull, blosum62, kind, dist, strip):
    # without statistics the global kernel leaves out the aligned-region word
    # wherever the distance does not read it (protein PercentSimilarity, the
    # word-mode distances); the distances are those of the oracle, and the
    # same as with statistics requested
    say("nw without statistics %s dist=%d strip=%d" % (kind, dist, strip))
    seqs, M, o, e = (protein_set, blosum62, 10, 1) if kind == "protein" else (dna_set, ednafull, 16, 4)
    nucleic = kind == "dna"
    al = ca.Aligner(M, o, e, mode=ca.MODE_NW_GLOBAL, distance_fn=dist,
                    alignment_type=ca.ALIGN_NUCLEIC if nucleic else ca.ALIGN_PROTEIN)
    al.set_schedule(ca.SCHED_PERSISTENT_I32, strip)
    al.set_sequences(seqs)
    n = len(seqs)
    r0, r1, c0, c1 = 0, 49, 50, n - 1
    got, _, pairs = al.compute_block(r0, r1, c0, c1, False)
    with_stats, _, _ = al.compute_block(r0, r1, c0, c1, False, stats=True)
    a, b = np.meshgrid(np.arange(r0, r1 + 1), np.arange(c0, c1 + 1), indexing="ij")
    want, _ = O.nw_pairs_mt(seqs.residues, seqs.offsets, a.ravel(), b.ravel(), M, o, e, nucleic=nucleic,
                            dist_fn=dist, threads=THREADS)
    assert pairs == a.size
    assert np.array_equal(got.ravel(), want), "%d distances differ" % (got.ravel() != want).sum()
    assert np.array_equal(got, with_stats)
    al.close()
