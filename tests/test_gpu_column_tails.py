"""GPU parity of the int16 score pass at every column-sequence tail.

The score pass reads a column sequence's codes four at a time, as whole
words, and clamps the word index to the sequence's trailing pad word, so
the columns past the sequence's end see the pad code without a per-column
length test.  These blocks put sequences of every length 0..40 (every
residue count mod 4, the empty sequence included) on the column side and on
the row side against longer ones, in the joint-profile path (DNA, <= 8
codes) and the per-code path (protein), in the score pass's max3 and plain
forms, and compare distances and statistics bit for bit with the CPU oracle.
"""
import random

import numpy as np
import pytest

import csharp_amd as ca
from tests.test_gpu_parity import assert_same, oracle_block, say

pytestmark = pytest.mark.gpu


def tails_set(alphabet, seed):
    rng = random.Random(seed)
    short = [bytes(rng.choice(alphabet) for _ in range(L)) for L in range(41)]
    base = bytes(rng.choice(alphabet) for _ in range(300))
    longer = []
    for k in range(24):
        s = bytearray(base[rng.randint(0, 60):rng.randint(200, 300)])
        for q in range(len(s)):
            if rng.random() < 0.1:
                s[q] = rng.choice(alphabet)
        longer.append(bytes(s))
    # short ones that share a prefix with the long ones, so their best
    # alignments run to their last residues
    short += [base[k:k + L] for k, L in ((3, 5), (10, 6), (20, 7), (0, 33), (40, 38))]
    return ca.SequenceSet.from_strings(short + longer), len(short)


@pytest.mark.parametrize("max3", [True, False])
@pytest.mark.parametrize("strip,score_strip", [(16, 40), (32, 24)])
def test_dna_joint_profile_tails(ednafull, strip, score_strip, max3):
    say("column tails dna strip=%d score strip=%d max3=%s" % (strip, score_strip, max3))
    seqs, ns = tails_set(b"ACGT", 31)
    n = len(seqs)
    al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT_I16)
    al.set_schedule(ca.SCHED_PERSISTENT_I16, strip)
    al.set_score_strip(score_strip)
    al.set_score_max3(max3)
    al.set_sequences(seqs)
    al.set_kernel_timing(True)
    for r0, r1, c0, c1 in ((ns, n - 1, 0, ns - 1), (0, ns - 1, ns, n - 1)):
        got, gst, _ = al.compute_block(r0, r1, c0, c1, False, stats=True)
        assert al.kernel_times()[0][0] == "score16"
        want, wst = oracle_block(seqs, r0, r1, c0, c1, False, ednafull, 16, 4)
        assert_same(got, want, gst, wst, "dna tails rows %d-%d cols %d-%d" % (r0, r1, c0, c1))
    al.close()


@pytest.mark.parametrize("max3", [True, False])
def test_protein_per_code_profile_tails(blosum62, max3):
    say("column tails protein max3=%s" % max3)
    seqs, ns = tails_set(b"ACDEFGHIKLMNPQRSTVWY", 32)
    n = len(seqs)
    al = ca.Aligner(blosum62, 10, 1, alignment_type=ca.ALIGN_PROTEIN, schedule=ca.SCHED_PERSISTENT_I16)
    al.set_score_max3(max3)
    al.set_sequences(seqs)
    al.set_kernel_timing(True)
    for r0, r1, c0, c1 in ((ns, n - 1, 0, ns - 1), (0, ns - 1, ns, n - 1)):
        got, gst, _ = al.compute_block(r0, r1, c0, c1, False, stats=True)
        assert al.kernel_times()[0][0] == "score16"
        want, wst = oracle_block(seqs, r0, r1, c0, c1, False, blosum62, 10, 1, nucleic=False)
        assert_same(got, want, gst, wst, "protein tails rows %d-%d cols %d-%d" % (r0, r1, c0, c1))
    al.close()
