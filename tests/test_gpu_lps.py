"""GPU parity of the int16 score pass with 16 or 32 lanes per pair
(lps16_kernel, SWG_SCHED_PERSISTENT_LPS16): the thin-block and serving score
pass, bit-exact against the CPU oracle through the C ABI -- stripes of 128
or 256 rows (queries of 1-2,000 residues), pre-columns and pad columns,
tiles of 8 or 4 column sequences with unequal lengths, saturating pairs, the
word-mode distances, the variant the library reports (and its choice of
lanes per pair) and its fallback for alphabets it cannot hold."""
import sys

import numpy as np
import pytest

import csharp_amd as ca
from oracle import oracle as O
from tests.test_gpu_parity import assert_same, oracle_block

pytestmark = pytest.mark.gpu

THREADS = O.cpu_threads()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("rows,ppw", [(3, 4), (64, 8)])
def test_thin_block_takes_the_lanes_per_pair_score_pass(ednafull, rows, ppw):
    """Blocks of 3 and 64 rows against 600 columns are thin: the autotuned
    schedule runs them with the int16 lanes-per-pair score pass, 32 lanes
    per pair (4 pairs per wave) where its 16-lane waves would not fill the
    GPU, else 16 (8 pairs per wave)."""
    say("lps16: thin block of %d rows, automatic choice" % rows)
    s = ca.synth_dna(31, rows + 600, 12, 0.10, 0.02, 0, [1, 900])
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(s)
    got, gst, _ = al.compute_block(0, rows - 1, rows, rows + 599, False, stats=True)
    cfg = al.last_block_config()
    assert cfg["variant"] == ca.SCHED_PERSISTENT_LPS16 and cfg["pairs_per_wave"] == ppw, cfg
    want, wst = oracle_block(s, 0, rows - 1, rows, rows + 599, False, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "thin block")
    al.close()


@pytest.mark.parametrize("lanes", [16, 32])
def test_queries_of_every_stripe_count(ednafull, lanes):
    """Row sequences of 1..2,000 residues (1 to 16 stripes of 128 rows or 1
    to 8 of 256, and each partial last stripe) against columns of 1..1,500
    residues sorted into tiles of 8 or 4 with very unequal members."""
    say("lps16: stripe counts, %d lanes per pair" % lanes)
    rng = np.random.default_rng(5)
    rows = [1, 2, 15, 16, 17, 127, 128, 129, 255, 256, 300, 641, 1000, 1537, 2000]
    fam = ca.synth_dna(32, 120, 6, 0.12, 0.03, 0, [1, 1500])
    seqs = [rng.choice(list(b"ACGT"), size=L).astype(np.uint8).tobytes() for L in rows]
    seqs[-3] = fam.seq(0)[:1000] + seqs[-3][1000:]  # a row related to a column
    s = ca.SequenceSet.from_strings(seqs + [fam.seq(i) for i in range(len(fam))])
    al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT_LPS16)
    al.set_lps_lanes(lanes)
    al.set_sequences(s)
    n = len(s)
    r1 = len(rows) - 1
    got, gst, _ = al.compute_block(0, r1, r1 + 1, n - 1, False, stats=True)
    cfg = al.last_block_config()
    assert cfg["variant"] == ca.SCHED_PERSISTENT_LPS16 and cfg["pairs_per_wave"] == 128 // lanes
    want, wst = oracle_block(s, 0, r1, r1 + 1, n - 1, False, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "stripe counts")
    al.close()


@pytest.mark.parametrize("lanes", [16, 32])
@pytest.mark.parametrize("o,e", [(16, 4), (5, 1), (4, 2)])
def test_low_complexity_and_ties(ednafull, o, e, lanes):
    """Repeats and homopolymers (many equal maxima across rows, columns and
    lanes), with penalties down to the smallest opening the byte table
    allows for EDNAFULL (4: mismatch -4 + 4 = 0)."""
    say("lps16: ties o=%d e=%d, %d lanes per pair" % (o, e, lanes))
    seqs = [b"A" * 300, b"A" * 37, b"AC" * 90, b"CA" * 77, b"ACGT" * 60, b"TGCA" * 41, b"AAAAC" * 50,
            b"G", b"C", b"GATTACA" * 30, b"T" * 5, b"ACGTTGCA" * 20]
    s = ca.SequenceSet.from_strings(seqs)
    al = ca.Aligner(ednafull, o, e, diag_mode=ca.DIAG_FULL, schedule=ca.SCHED_PERSISTENT_LPS16)
    al.set_lps_lanes(lanes)
    al.set_sequences(s)
    n = len(s)
    got, gst, _ = al.compute_block(0, n - 1, 0, n - 1, True, stats=True)
    assert al.last_block_config()["variant"] == ca.SCHED_PERSISTENT_LPS16
    want, wst = oracle_block(s, 0, n - 1, 0, n - 1, False, ednafull, o, e)
    assert_same(got, want, gst, wst, "ties")
    al.close()


@pytest.mark.parametrize("lanes", [16, 32])
def test_saturating_and_empty_pairs(ednafull, lanes):
    """Related 7,000-residue pairs score above the int16 range: flagged for
    the whole-matrix int32 re-run; empty sequences flagged too."""
    say("lps16: saturation")
    rng = np.random.default_rng(9)
    anc = rng.choice(list(b"ACGT"), size=7200).astype(np.uint8)
    a = anc[:7000].tobytes()
    b = anc[100:7200].copy()
    b[::50] = ord("A")
    s = ca.SequenceSet.from_strings([a, b"", b.tobytes(), b"ACGT" * 30, b""])
    al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT_LPS16)
    al.set_lps_lanes(lanes)
    al.set_sequences(s)
    got, gst, _ = al.compute_block(0, 1, 2, 4, False, stats=True)
    want, wst = oracle_block(s, 0, 1, 2, 4, False, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "saturation")
    assert wst[0, 0, 0] > 32766
    al.close()


@pytest.mark.parametrize("dist,tdist", [(ca.DIST_PERCENT_IDENTITY, ca.TDIST_DOUBLE), (ca.DIST_KIMURA2, ca.TDIST_INT16),
                                        (ca.DIST_JUKES_CANTOR, ca.TDIST_DOUBLE)])
def test_distances(ednafull, dist, tdist):
    """Double PercentIdentity bit-exact; Kimura2 / Jukes-Cantor within the
    4-ulp tolerance of the logarithm for Double (bit-exact Int16)."""
    say("lps16: distance %d tdist %d" % (dist, tdist))
    s = ca.synth_dna(33, 300, 8, 0.10, 0.02, 0, [20, 700])
    al = ca.Aligner(ednafull, 16, 4, distance_fn=dist, tdistance=tdist, schedule=ca.SCHED_PERSISTENT_LPS16)
    al.set_sequences(s)
    got, _, _ = al.compute_block(0, 9, 10, 299, False, stats=True)
    assert al.last_block_config()["variant"] == ca.SCHED_PERSISTENT_LPS16
    a, b = np.meshgrid(np.arange(0, 10), np.arange(10, 300), indexing="ij")
    want, _ = O.sw_pairs_mt(s.residues, s.offsets, a.ravel(), b.ravel(), ednafull, 16, 4, dist_fn=dist,
                            tdist_double=tdist == ca.TDIST_DOUBLE, threads=THREADS)
    want = want.reshape(got.shape)
    nan = np.isnan(want) if want.dtype == np.float64 else np.zeros(want.shape, bool)
    if want.dtype == np.float64:
        assert np.array_equal(np.isnan(got), nan)
    if tdist == ca.TDIST_DOUBLE and dist != ca.DIST_PERCENT_IDENTITY:
        g, w = got[~nan], want[~nan]
        assert (np.abs(g - w) <= np.abs(w) * 2.0 ** -50).all()  # 4 ulp: the logarithm
    else:
        assert np.array_equal(got[~nan], want[~nan])
    al.close()


def test_large_alphabet_falls_back(blosum62):
    """Protein (25 codes) cannot use the byte table: the forced schedule runs
    the int32 sixteen-lanes-per-pair kernel instead."""
    say("lps16: protein fallback")
    s = ca.synth_protein(34, 60, 2, [120, 0.5, 1, 400])
    al = ca.Aligner(blosum62, 10, 1, alignment_type=ca.ALIGN_PROTEIN, schedule=ca.SCHED_PERSISTENT_LPS16)
    al.set_sequences(s)
    got, gst, _ = al.compute_block(0, 3, 4, 59, False, stats=True)
    assert al.last_block_config()["variant"] == ca.SCHED_PERSISTENT_LPP16
    want, wst = oracle_block(s, 0, 3, 4, 59, False, blosum62, 10, 1, nucleic=False)
    assert_same(got, want, gst, wst, "protein fallback")
    al.close()


def test_repeated_blocks_of_other_shapes(ednafull):
    """One context, several thin blocks of different shapes in a row (the
    captured graph is replayed and re-captured), each against the oracle."""
    say("lps16: repeated blocks")
    s = ca.synth_dna(35, 400, 10, 0.10, 0.02, 0, [1, 600])
    al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT_LPS16)
    al.set_sequences(s)
    for k, (r0, r1, c0, c1) in enumerate([(0, 4, 5, 399), (10, 11, 100, 399), (20, 27, 30, 90), (0, 4, 5, 399)]):
        al.set_lps_lanes((0, 16, 32, 16)[k])
        got, gst, _ = al.compute_block(r0, r1, c0, c1, False, stats=True)
        want, wst = oracle_block(s, r0, r1, c0, c1, False, ednafull, 16, 4)
        assert_same(got, want, gst, wst, "block %d-%d x %d-%d" % (r0, r1, c0, c1))
    al.close()
