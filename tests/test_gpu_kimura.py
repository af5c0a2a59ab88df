"""GPU parity of the Kimura2 and Jukes-Cantor distances (through the C ABI)
against the CPU oracle.

ComputeDistance (Salsa.SmithWatermanMPI/Program.cs:259-314) with
PairwiseAlignment.ComputeKimuraDistance / ComputeJukesCantorDistance
(Salsa.CoreTPL/Bio/Algorithms/Alignment.cs:254-348), Int16 and Double, over
every schedule.  The counts behind them (aligned columns without gap columns,
transitions, transversions, differences) and the Int16 distances must be
bit-exact.  The Double distances go through the natural logarithm: the GPU
uses the device library's log, the oracle the host C library's, and neither
is correctly rounded, so a Double distance may differ from the oracle's by a
few units in the last place: tolerance |got - want| <= 2^-50 |want| (4 ulp).
"""
import sys

import numpy as np
import pytest

import csharp_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = O.cpu_threads()
DISTS = {"kimura2": ca.DIST_KIMURA2, "jukes_cantor": ca.DIST_JUKES_CANTOR}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_block(seqs, r0, r1, c0, c1, diag_lower, M, o, e, dist_fn, tdouble):
    rows, cols = r1 - r0 + 1, c1 - c0 + 1
    a, b = [], []
    for i in range(r0, r1 + 1):
        for j in range(c0, c1 + 1):
            if diag_lower and j >= i:
                continue
            a.append(i); b.append(j)
    d, st = O.sw_pairs_mt(seqs.residues, seqs.offsets, a, b, M, o, e, nucleic=True, dist_fn=dist_fn,
                          tdist_double=tdouble, threads=THREADS, want_stats=True)
    out = np.zeros((rows, cols), dtype=np.float64 if tdouble else np.int16)
    stats = np.zeros((rows, cols, 4), dtype=np.int32)
    a = np.asarray(a) - r0
    b = np.asarray(b) - c0
    out[a, b] = d
    stats[a, b] = st
    if diag_lower:
        out[b, a] = d
        stats[b, a] = st
    return out, stats


def assert_match(got, want, gst, wst, what):
    assert np.array_equal(gst, wst), "%s: statistics differ at %s" % (what, np.argwhere(gst != wst)[:5].tolist())
    if got.dtype == np.int16:
        bad = np.argwhere(got != want)
        assert bad.size == 0, "%s: %d Int16 mismatches, first %s got %s want %s stats %s" % (
            what, len(bad), bad[:3].tolist(), [got[tuple(i)] for i in bad[:3]], [want[tuple(i)] for i in bad[:3]],
            [gst[tuple(i)].tolist() for i in bad[:3]])
        return
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan), what + ": NaN positions differ"
    g, w = got[~nan], want[~nan]
    ok = np.abs(g - w) <= np.abs(w) * 2.0 ** -50
    assert ok.all(), "%s: %d Double distances outside 4 ulp, first got %r want %r" % (
        what, (~ok).sum(), g[~ok][:3].tolist(), w[~ok][:3].tolist())
    say("%s: %d of %d Double distances bit-identical" % (what, int((g == w).sum()), g.size))


@pytest.fixture(scope="module")
def dna_sets():
    """Per distance: related and unrelated DNA plus edge sequences.  The
    Jukes-Cantor set leaves out U: EDNAFULL scores U/T +5 as a difference, so
    an alignment of U against T alone is a local optimum with d < 0, where
    the reference throws (that case has its own test)."""
    s = ca.synth_dna(11, 180, 12, 0.10, 0.02, 0, [1, 260])
    extra = [b"", b"A", b"ACGT" * 40, b"A" * 150, b"ACACACACAC" * 12, b"NNNNACGTNN", b"ACGUXACGT", b"T" * 300,
             b"AGAGAGAGAGAG" * 5, b"GAGAGAGAGAGA" * 5, b"CTCTCTCT" * 9]
    base = [s.seq(i) for i in range(len(s))]
    jc = [b"ACGTXACGT" if x == b"ACGUXACGT" else x for x in extra]
    return {"kimura2": ca.SequenceSet.from_strings(base + extra),
            "jukes_cantor": ca.SequenceSet.from_strings(base + jc)}


SCHEDULES = [ca.SCHED_PERSISTENT, ca.SCHED_PLAIN, ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16,
             ca.SCHED_PERSISTENT_LPP16]


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("tdouble", [False, True])
@pytest.mark.parametrize("dist", sorted(DISTS))
def test_offdiag_block(dna_sets, ednafull, dist, tdouble, schedule):
    say("%s offdiag double=%d schedule=%d" % (dist, tdouble, schedule))
    dna_set = dna_sets[dist]
    al = ca.Aligner(ednafull, 16, 4, distance_fn=DISTS[dist],
                    tdistance=ca.TDIST_DOUBLE if tdouble else ca.TDIST_INT16, schedule=schedule)
    al.set_sequences(dna_set)
    n = len(dna_set)
    r0, r1, c0, c1 = 0, 89, 90, n - 1
    got, gst, pairs = al.compute_block(r0, r1, c0, c1, False, stats=True)
    want, wst = oracle_block(dna_set, r0, r1, c0, c1, False, ednafull, 16, 4, DISTS[dist], tdouble)
    assert pairs == (r1 - r0 + 1) * (c1 - c0 + 1)
    assert_match(got, want, gst, wst, dist)
    al.close()


@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16])
@pytest.mark.parametrize("dist", sorted(DISTS))
def test_diagonal_block_lower_mirror(dna_sets, ednafull, dist, schedule):
    say("%s diag schedule=%d" % (dist, schedule))
    dna_set = dna_sets[dist]
    al = ca.Aligner(ednafull, 16, 4, distance_fn=DISTS[dist], schedule=schedule)
    al.set_sequences(dna_set)
    n = len(dna_set)
    r0, r1 = 60, n - 1
    got, gst, _ = al.compute_block(r0, r1, r0, r1, True, stats=True)
    want, wst = oracle_block(dna_set, r0, r1, r0, r1, True, ednafull, 16, 4, DISTS[dist], False)
    assert_match(got, want, gst, wst, dist + " diag")
    assert np.all(np.diag(got) == 0)
    al.close()


def test_kimura_big_rectangles(ednafull):
    """Related long pairs: the two-pass schedule's one-wave-per-pair re-run
    runs both Kimura2 passes."""
    say("kimura2 big rectangles")
    from tests.test_gpu_parity import _mutated_family
    rng = np.random.default_rng(41)
    s = _mutated_family(rng, b"ACGT", 24, 600, 1300, 0.10, 0.03)
    al = ca.Aligner(ednafull, 16, 4, distance_fn=ca.DIST_KIMURA2, schedule=ca.SCHED_PERSISTENT_I16)
    al.set_sequences(s)
    n = len(s)
    al.set_kernel_timing(True)
    got, gst, _ = al.compute_block(0, 23, 0, n - 1, False, stats=True)
    names = [k[0] for k in al.kernel_times()]
    al.set_kernel_timing(False)
    assert names.count("rect64") == 2 and names.count("rerun_capture") == 2
    want, wst = oracle_block(s, 0, 23, 0, n - 1, False, ednafull, 16, 4, ca.DIST_KIMURA2, False)
    assert_match(got, want, gst, wst, "kimura2 big")
    got2, gst2, _ = al.compute_block(0, 23, 0, n - 1, False, stats=True)  # graph path
    assert np.array_equal(got2, got) and np.array_equal(gst2, gst)
    al.close()


def positive_mismatch_matrix():
    M = np.zeros((127, 127), dtype=np.float32)
    for a in b"ACGT":
        for b in b"ACGT":
            M[a, b] = 2 if a == b else 1
    return M


@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16])
def test_jukes_cantor_undefined_raises(ednafull, schedule):
    """Mismatches scoring above zero make all-different alignments: the
    reference throws; the block reports SWG_E_JC_UNDEFINED, and the context
    stays usable."""
    say("jukes-cantor undefined schedule=%d" % schedule)
    M = positive_mismatch_matrix()
    seqs = ca.SequenceSet.from_strings([b"AAAAAAAA", b"CCCCCCCC", b"ACGTACGT", b"ACGTACGA"])
    al = ca.Aligner(M, 16, 4, distance_fn=ca.DIST_JUKES_CANTOR, schedule=schedule)
    al.set_sequences(seqs)
    with pytest.raises(ca.JukesCantorUndefined) as ei:
        al.compute_block(0, 0, 1, 1, False)
    assert ei.value.code == ca.SWG_E_JC_UNDEFINED
    # the oracle agrees: U/T scores +5 under EDNAFULL and is one difference
    ut = ca.SequenceSet.from_strings([b"T" * 30, b"ACGUXACGT"])
    with pytest.raises(ArithmeticError):
        O.sw_pairs_mt(ut.residues, ut.offsets, [0], [1], ednafull, 16, 4, dist_fn=ca.DIST_JUKES_CANTOR)
    al2 = ca.Aligner(ednafull, 16, 4, distance_fn=ca.DIST_JUKES_CANTOR, schedule=schedule)
    al2.set_sequences(ut)
    with pytest.raises(ca.JukesCantorUndefined):
        al2.compute_block(0, 0, 1, 1, False)
    al2.close()
    got, gst, _ = al.compute_block(2, 2, 3, 3, False, stats=True)  # one difference in 8 columns
    assert gst[0, 0, 1:].tolist() == [8, 1, 0]
    assert got[0, 0] == int(O.jukes_cantor(8, 0, 1) * 32767)
    al.close()


def test_kimura_artificial_distance_wraps():
    """1 - (2P + Q) <= 0 gives the artificial distance 10; in Int16,
    (short)(10 * 32767) keeps the low 16 bits of 327670: -10."""
    say("kimura2 artificial distance")
    M = positive_mismatch_matrix()
    seqs = ca.SequenceSet.from_strings([b"AAAAAAAA", b"GGGGGGGG"])
    for td, want in ((ca.TDIST_INT16, -10), (ca.TDIST_DOUBLE, 10.0)):
        al = ca.Aligner(M, 16, 4, distance_fn=ca.DIST_KIMURA2, tdistance=td, schedule=ca.SCHED_PERSISTENT_I32)
        al.set_sequences(seqs)
        got, gst, _ = al.compute_block(0, 0, 1, 1, False, stats=True)
        assert got[0, 0] == want and gst[0, 0, 1:].tolist() == [8, 8, 0]
        al.close()


def test_gap_character_is_refused(ednafull):
    say("kimura2 '-' residue")
    al = ca.Aligner(ednafull, 16, 4, distance_fn=ca.DIST_KIMURA2)
    with pytest.raises(ca.SwgError) as ei:
        al.set_sequences(ca.SequenceSet.from_strings([b"ACGT", b"AC-T"]))
    assert ei.value.code == ca.SWG_E_UNSUPPORTED
    al.close()


def test_serving_kimura_and_undefined_query(ednafull):
    """A server with Kimura2 matches the oracle; with Jukes-Cantor a query
    whose alignment is undefined fails alone, its batch mates complete."""
    say("serving kimura2 / jukes-cantor")
    panel = ca.synth_dna(21, 200, 10, 0.10, 0.02, 0, [30, 250])
    qs = ca.synth_dna(22, 6, 10, 0.10, 0.02, 0, [1, 300])
    queries = [qs.seq(i) for i in range(len(qs))]
    srv = ca.Server(panel, ednafull, 16, 4, distance_fn=ca.DIST_KIMURA2, max_query_len=400, max_batch=4, lanes=2,
                    hbm_budget=64 << 20)
    tickets = [srv.submit(q) for q in queries]
    allseq = ca.SequenceSet.from_strings([panel.seq(i) for i in range(len(panel))] + queries)
    npan = len(panel)
    for k, t in enumerate(tickets):
        assert srv.wait(t, 60000)
        out, st = srv.result(t)
        d, wst = O.sw_pairs_mt(allseq.residues, allseq.offsets, np.full(npan, npan + k), np.arange(npan), ednafull,
                               16, 4, dist_fn=ca.DIST_KIMURA2, threads=THREADS, want_stats=True)
        assert np.array_equal(out, d) and np.array_equal(st, wst), "query %d" % k
    srv.close()

    M = positive_mismatch_matrix()
    panel = ca.SequenceSet.from_strings([b"CCCCCCCC", b"CCCCGGGG"])
    srv = ca.Server(panel, M, 16, 4, distance_fn=ca.DIST_JUKES_CANTOR, max_query_len=16, max_batch=4, lanes=1,
                    hbm_budget=1 << 20)
    t_bad = srv.submit(b"AAAAAAAA")      # 8 differences in 8 columns
    t_good = srv.submit(b"CCCCGGGA")     # 4 of 8 and 1 of 8
    with pytest.raises(ca.SwgError) as ei:
        srv.wait(t_bad, 60000)
    assert ei.value.code == ca.SWG_E_JC_UNDEFINED
    assert srv.wait(t_good, 60000)
    out, st = srv.result(t_good)
    assert st[1, 1:].tolist() == [8, 1, 0] and out[1] == int(O.jukes_cantor(8, 0, 1) * 32767)
    srv.close()
