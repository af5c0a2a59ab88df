"""GPU parity of the score pass's max3 form (values biased by gap_open, the
three-way maxima as v_pk_maximum3_f16 on int16 bit patterns, the row
maximum over two columns at once) against its plain form and the CPU oracle
through the C ABI: joint and per-code profiles, every strip height, scores
straddling the form's range limit (0x7C00 - gap_open: those pairs are
flagged for the int32 re-run over their whole matrix) and the int16
saturation, and penalties the form cannot take (score + gap_open < 0), which
run the plain form."""
import sys

import numpy as np
import pytest

import csharp_amd as ca
from oracle import oracle as O
from tests.test_gpu_parity import assert_same, oracle_block

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(s, M, o, e, blk, max3, strip=0, stats=True, **kw):
    al = ca.Aligner(M, o, e, schedule=ca.SCHED_PERSISTENT_I16, **kw)
    al.set_score_max3(max3)
    if strip:
        al.set_score_strip(strip)
    al.set_sequences(s)
    got, gst, _ = al.compute_block(*blk, False, stats=stats)
    assert al.last_block_config()["variant"] == ca.SCHED_PERSISTENT_I16
    al.close()
    return got, gst


@pytest.mark.parametrize("strip", [16, 24, 32, 40])
def test_max3_form_matches_the_plain_form_and_the_oracle(ednafull, strip):
    """Related and unrelated DNA of 1..1,500 residues (tiles of very unequal
    members, pad rows in every last strip)."""
    say("max3 form vs plain, strip %d" % strip)
    s = ca.synth_dna(71, 700, 10, 0.12, 0.03, 0, [1, 1500])
    blk = (0, 199, 200, 699)
    got3, gst3 = run(s, ednafull, 16, 4, blk, True, strip)
    got0, gst0 = run(s, ednafull, 16, 4, blk, False, strip)
    assert np.array_equal(got3, got0) and np.array_equal(gst3, gst0)
    rng = np.random.default_rng(strip)
    a = rng.integers(0, 200, 3000)
    b = rng.integers(200, 700, 3000)
    d, st = O.sw_pairs_mt(s.residues, s.offsets, a, b, ednafull, 16, 4, threads=O.cpu_threads(),
                          want_stats=True)
    assert np.array_equal(got3[a, b - 200], d)
    assert np.array_equal(gst3[a, b - 200][:, :3], st[:, :3])


def test_scores_straddling_the_range_limit(ednafull):
    """Prefixes of one 6,700-residue sequence: pair scores 5 x the shorter
    length, 31,700 .. 33,500 -- below, at and past the max3 form's limit
    (0x7C00 - 16 = 31,728) and the int16 saturation (32,766)."""
    say("max3 form: range limit")
    rng = np.random.default_rng(5)
    x = rng.choice(list(b"ACGT"), size=6700).astype(np.uint8).tobytes()
    lens = [6340, 6345, 6346, 6347, 6360, 6553, 6700]
    s = ca.SequenceSet.from_strings([x[:n] for n in lens] + [x[:n] for n in lens])
    k = len(lens)
    blk = (0, k - 1, k, 2 * k - 1)
    want, wst = oracle_block(s, *blk, False, ednafull, 16, 4)
    for max3 in (True, False):
        got, gst = run(s, ednafull, 16, 4, blk, max3)
        assert_same(got, want, gst, wst, "range limit, max3 %s" % max3)
    assert (wst[..., 0] == 5 * np.minimum.outer(lens, lens)).all()


def test_protein_per_code_profile(blosum62):
    """25 residue codes: the per-code profile (one v_perm per row) in the max3
    form (BLOSUM62 10/1: score + gap_open >= 6)."""
    say("max3 form: protein")
    s = ca.synth_protein(72, 300, 2, [200, 0.5, 1, 600])
    blk = (0, 39, 40, 299)
    want, wst = oracle_block(s, *blk, False, blosum62, 10, 1, nucleic=False)
    for max3 in (True, False):
        got, gst = run(s, blosum62, 10, 1, blk, max3, alignment_type=ca.ALIGN_PROTEIN)
        assert_same(got, want, gst, wst, "protein, max3 %s" % max3)


@pytest.mark.parametrize("o,e", [(2, 1), (3, 3), (0, 0)])
def test_penalties_outside_the_form_run_the_plain_form(ednafull, o, e):
    """With gap_open < 4 a mismatch (-4) plus gap_open is negative: the set
    cannot run the max3 form and the library runs the plain one."""
    say("max3 form: o=%d e=%d" % (o, e))
    s = ca.synth_dna(73, 160, 6, 0.12, 0.03, 0, [1, 300])
    blk = (0, 59, 60, 159)
    want, wst = oracle_block(s, *blk, False, ednafull, o, e)
    got, gst = run(s, ednafull, o, e, blk, True)
    assert_same(got, want, gst, wst, "o=%d e=%d" % (o, e))


def test_open_zero_and_extension_above_opening(ednafull):
    """A matrix with no negative score (match 3, mismatch 0) takes the max3
    form even at gap_open 0 (pads then tie real cells, which come first in
    row-major order), and with gap_ext above gap_open."""
    say("max3 form: gap_open 0")
    M = np.where(np.eye(ednafull.shape[0], dtype=bool), 3, 0).astype(ednafull.dtype)
    s = ca.synth_dna(74, 200, 5, 0.15, 0.04, 0, [1, 400])
    blk = (0, 69, 70, 199)
    for o, e in ((0, 0), (0, 2), (1, 3)):
        want, wst = oracle_block(s, *blk, False, M, o, e)
        for max3 in (True, False):
            got, gst = run(s, M, o, e, blk, max3)
            assert_same(got, want, gst, wst, "o=%d e=%d max3 %s" % (o, e, max3))


def test_reverse_row_cap_outside_its_biased_form(ednafull):
    """A reverse row cap of 4,096 rows lets a clamped value climb 4,096 x 5
    > 16,384: the reverse pass then runs its plain form while the score pass
    keeps its max3 form; both caps give the oracle's distances."""
    say("max3 form: reverse row cap 4096")
    s = ca.synth_dna(75, 300, 3, 0.08, 0.02, 0, [600, 2500])
    blk = (0, 49, 50, 299)
    want, wst = oracle_block(s, *blk, False, ednafull, 16, 4)
    for cap in (4096, 512):
        al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT_I16)
        al.set_reverse_row_cap(cap)
        al.set_sequences(s)
        got, gst, _ = al.compute_block(*blk, False, stats=True)
        al.close()
        assert_same(got, want, gst, wst, "row cap %d" % cap)
