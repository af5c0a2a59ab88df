"""GPU parity of the SWMS and NeedlemanWunsch ComputeDistance helpers
(through the C ABI) against the CPU oracle.

SWG_PROGRAM_SWMS selects Salsa.SmithWatermanMS/Program.cs:417-668 on the
local alignment: PercentSimilarity over non-gap columns scoring > 0,
Kimura2 clamped to 10, Jukes-Cantor 10 where undefined, MinMaxNormScore with
minimum 0.  The global mode follows Salsa.NeedlemanWunschTPL/Program.cs:
MinMaxNormScore with minimum 2*open + (|si|+|sj|)*ext.  Bit-exact for the
Int16 distances and the statistics.  (Both programs align with the closed
Bio.dll; the alignment here is this build's, so the reference-level parity
of these programs stays unpinned.)
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


def expected(seqs, r0, r1, c0, c1, M, o, e, **kw):
    a, b = np.meshgrid(np.arange(r0, r1 + 1), np.arange(c0, c1 + 1), indexing="ij")
    d, st = O.sw_pairs_mt(seqs.residues, seqs.offsets, a.ravel(), b.ravel(), M, o, e, threads=THREADS,
                          want_stats=True, **kw)
    return d.reshape(a.shape), st.reshape(a.shape + (4,))


@pytest.fixture(scope="module")
def dna_set():
    s = ca.synth_dna(13, 150, 10, 0.10, 0.02, 0, [1, 260])
    extra = [b"", b"A", b"ACGT" * 30, b"T" * 200, b"NNNNACGTNN", b"ACGUXACGT"]
    return ca.SequenceSet.from_strings([s.seq(i) for i in range(len(s))] + extra)


@pytest.fixture(scope="module")
def protein_set():
    s = ca.synth_protein(12, 100, 2, [120, 0.5, 1, 400])
    extra = [b"", b"W", b"XXBZ*AR", b"WWXWW", b"WWXXWW", b"ACDEFGHIKLMNPQRSTVWYX" * 4]
    return ca.SequenceSet.from_strings([s.seq(i) for i in range(len(s))] + extra)


@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT_LPP16])
def test_swms_protein_similarity(protein_set, blosum62, schedule):
    say("swms protein similarity schedule=%d" % schedule)
    al = ca.Aligner(blosum62, 10, 1, alignment_type=ca.ALIGN_PROTEIN, schedule=schedule, program=ca.PROGRAM_SWMS)
    al.set_sequences(protein_set)
    n = len(protein_set)
    got, gst, _ = al.compute_block(0, 50, 51, n - 1, False, stats=True)
    want, wst = expected(protein_set, 0, 50, 51, n - 1, blosum62, 10, 1, nucleic=False, program=1)
    assert np.array_equal(gst[..., :2], wst[..., :2])
    assert np.array_equal(gst[..., 2], wst[..., 3]), "positive-score column counts differ"
    assert np.array_equal(got, want)
    # the edge sequences (X/X scores -1) against every sequence: the SWMPI rule
    # counts identical pairs even where they score <= 0, this one does not
    got, gst, _ = al.compute_block(n - 6, n - 1, 0, n - 1, False, stats=True)
    want, wst = expected(protein_set, n - 6, n - 1, 0, n - 1, blosum62, 10, 1, nucleic=False, program=1)
    assert np.array_equal(gst[..., 2], wst[..., 3]) and np.array_equal(got, want)
    _, st_mpi = expected(protein_set, n - 6, n - 1, 0, n - 1, blosum62, 10, 1, nucleic=False)
    assert (st_mpi[..., 3] > wst[..., 3]).any()
    al.close()


@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT, ca.SCHED_PLAIN, ca.SCHED_PERSISTENT_I32,
                                      ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT_LPP16])
def test_swms_minmax(dna_set, ednafull, schedule):
    say("swms minmax schedule=%d" % schedule)
    al = ca.Aligner(ednafull, 16, 4, distance_fn=ca.DIST_MINMAX_NORM, schedule=schedule, program=ca.PROGRAM_SWMS,
                    diag_mode=ca.DIAG_FULL)
    al.set_sequences(dna_set)
    n = len(dna_set)
    got, gst, _ = al.compute_block(0, 60, 61, n - 1, False, stats=True)
    want, wst = expected(dna_set, 0, 60, 61, n - 1, ednafull, 16, 4, dist_fn=3, program=1)
    assert np.array_equal(gst[..., :3], wst[..., :3])
    assert np.array_equal(got, want)
    # full diagonal block (SWMS computes every (i, j), i == j included)
    got, gst, _ = al.compute_block(40, 80, 40, 80, True, stats=True)
    want, wst = expected(dna_set, 40, 80, 40, 80, ednafull, 16, 4, dist_fn=3, program=1)
    assert np.array_equal(got, want)
    al.close()


def test_swms_minmax_double_is_zero(dna_set, ednafull):
    say("swms minmax double")
    al = ca.Aligner(ednafull, 16, 4, distance_fn=ca.DIST_MINMAX_NORM, tdistance=ca.TDIST_DOUBLE,
                    program=ca.PROGRAM_SWMS)
    al.set_sequences(dna_set)
    got, _, _ = al.compute_block(0, 20, 21, 60, False)
    assert np.all(got == 0.0)
    al.close()


@pytest.mark.parametrize("dist", [ca.DIST_KIMURA2, ca.DIST_JUKES_CANTOR, ca.DIST_PERCENT_IDENTITY])
def test_swms_nucleic_distances(dna_set, ednafull, dist):
    say("swms nucleic dist=%d" % dist)
    sub = dna_set.subset([i for i in range(len(dna_set)) if dna_set.seq(i) != b"ACGUXACGT"])
    al = ca.Aligner(ednafull, 16, 4, distance_fn=dist, schedule=ca.SCHED_PERSISTENT_I16, program=ca.PROGRAM_SWMS)
    al.set_sequences(sub)
    n = len(sub)
    got, gst, _ = al.compute_block(0, 50, 51, n - 1, False, stats=True)
    want, wst = expected(sub, 0, 50, 51, n - 1, ednafull, 16, 4, dist_fn=dist, program=1)
    assert np.array_equal(gst, wst) if dist != ca.DIST_PERCENT_IDENTITY else \
        np.array_equal(gst[..., :3], wst[..., :3])
    assert np.array_equal(got, want)
    al.close()


def test_swms_jukes_cantor_undefined_is_ten():
    say("swms jukes-cantor 10")
    M = np.zeros((127, 127), dtype=np.float32)
    for a in b"ACGT":
        for b in b"ACGT":
            M[a, b] = 2 if a == b else 1
    seqs = ca.SequenceSet.from_strings([b"AAAAAAAA", b"CCCCCCCC"])
    for program, raises in ((ca.PROGRAM_SWMPI, True), (ca.PROGRAM_SWMS, False)):
        al = ca.Aligner(M, 16, 4, distance_fn=ca.DIST_JUKES_CANTOR, schedule=ca.SCHED_PERSISTENT_I32,
                        program=program)
        al.set_sequences(seqs)
        if raises:
            with pytest.raises(ca.JukesCantorUndefined):
                al.compute_block(0, 0, 1, 1, False)
        else:
            got, _, _ = al.compute_block(0, 0, 1, 1, False)
            assert got[0, 0] == -10  # (short)(10 * 32767)
        al.close()


def test_global_minmax(dna_set, ednafull):
    say("global minmax")
    al = ca.Aligner(ednafull, 16, 4, distance_fn=ca.DIST_MINMAX_NORM, mode=ca.MODE_NW_GLOBAL)
    al.set_sequences(dna_set)
    n = len(dna_set)
    got, gst, _ = al.compute_block(0, 50, 51, n - 1, False, stats=True)
    a, b = np.meshgrid(np.arange(0, 51), np.arange(51, n), indexing="ij")
    want, wst = O.nw_pairs_mt(dna_set.residues, dna_set.offsets, a.ravel(), b.ravel(), ednafull, 16, 4,
                              dist_fn=3, threads=THREADS, want_stats=True)
    assert np.array_equal(gst.reshape(-1, 4)[:, 0], wst[:, 0])
    assert np.array_equal(got.ravel(), want)
    al.close()


@pytest.mark.parametrize("tdouble", [False, True])
@pytest.mark.parametrize("dist", [ca.DIST_KIMURA2, ca.DIST_JUKES_CANTOR])
def test_global_kimura_jukes_cantor(dna_set, ednafull, dist, tdouble):
    """The global kernel carries the Kimura2 / Jukes-Cantor counts like the
    local ones (Kimura2: two passes); the global program's helpers clamp
    Kimura2 at 10 and give Jukes-Cantor 10 where undefined.  Int16 and the
    statistics bit-exact, Double within 4 ulp (the logarithm)."""
    say("global dist=%d double=%d" % (dist, tdouble))
    al = ca.Aligner(ednafull, 16, 4, distance_fn=dist, mode=ca.MODE_NW_GLOBAL,
                    tdistance=ca.TDIST_DOUBLE if tdouble else ca.TDIST_INT16)
    al.set_sequences(dna_set)
    n = len(dna_set)
    got, gst, _ = al.compute_block(0, 50, 51, n - 1, False, stats=True)
    a, b = np.meshgrid(np.arange(0, 51), np.arange(51, n), indexing="ij")
    want, wst = O.nw_pairs_mt(dna_set.residues, dna_set.offsets, a.ravel(), b.ravel(), ednafull, 16, 4,
                              dist_fn=dist, tdist_double=tdouble, threads=THREADS, want_stats=True)
    assert np.array_equal(gst.reshape(-1, 4), wst)
    g = got.ravel()
    if tdouble:
        nan = np.isnan(want)
        assert np.array_equal(np.isnan(g), nan)
        assert np.all(np.abs(g[~nan] - want[~nan]) <= np.abs(want[~nan]) * 2.0 ** -50)
    else:
        assert np.array_equal(g, want)
    al.close()


def test_minmax_needs_swms_or_global(ednafull):
    with pytest.raises(ca.SwgError) as ei:
        ca.Aligner(ednafull, 16, 4, distance_fn=ca.DIST_MINMAX_NORM)
    assert ei.value.code == ca.SWG_E_UNSUPPORTED
