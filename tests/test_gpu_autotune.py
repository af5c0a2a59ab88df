"""The autotuner's search space, on the GPU through the C ABI: every
configuration it can choose (variant x tile shape) is bit-exact against the
oracle when forced, and an autotuned run measures the variants, then the
winner's strips, each trial on real work whose results are kept, and settles
on one of the configurations it measured."""
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


@pytest.fixture(scope="module")
def wide_set():
    # 3,000 row sequences x 256 columns: 12,000 waves one lane per pair, not
    # a thin block, so the autotuner (not the 16-lane shortcut) decides
    return ca.synth_dna(61, 3256, 20, 0.10, 0.02, 0, [30, 90])


@pytest.fixture(scope="module")
def wide_expected(wide_set, ednafull):
    a, b = np.meshgrid(np.arange(0, 3000), np.arange(3000, 3256), indexing="ij")
    d, _ = O.sw_pairs_mt(wide_set.residues, wide_set.offsets, a.ravel(), b.ravel(), ednafull, 16, 4,
                         threads=THREADS)
    return d.reshape(a.shape)


@pytest.mark.parametrize("score_strip", [16, 24, 32, 40])
@pytest.mark.parametrize("rerun_strip", [16, 32])
def test_forced_two_pass_tile_shapes(wide_set, wide_expected, ednafull, score_strip, rerun_strip):
    say("two-pass score strip %d re-run strip %d" % (score_strip, rerun_strip))
    al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT_I16)
    al.set_sequences(wide_set)
    al.set_schedule(ca.SCHED_PERSISTENT_I16, rerun_strip)
    al.set_score_strip(score_strip)
    got, _, _ = al.compute_block(0, 2999, 3000, 3255, False)
    cfg = al.last_block_config()
    assert (cfg["variant"], cfg["strip_rows"], cfg["rerun_strip_rows"], cfg["pairs_per_wave"], cfg["tuning"]) == (
        ca.SCHED_PERSISTENT_I16, score_strip, rerun_strip, 128, 0)
    assert np.array_equal(got, wide_expected)
    al.close()


def run_until_tuned(al, blk, expected, limit=10):
    trials = []
    for k in range(limit):
        got, _, _ = al.compute_block(*blk, False)
        cfg = al.last_block_config()
        assert np.array_equal(got, expected), "block %d (%s) differs from the oracle" % (k, cfg)
        say("block %d: %s" % (k, cfg))
        if cfg["tuning"] == 2:
            return trials, cfg
        assert cfg["tuning"] == 1
        trials.append(cfg)
    raise AssertionError("the autotuner did not settle in %d blocks: %s" % (limit, trials))


def shape(cfg):
    return cfg["variant"], cfg["strip_rows"], cfg["rerun_strip_rows"]


def test_autotuner_searches_variant_then_tile_shape(wide_set, wide_expected, ednafull):
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(wide_set)
    al.set_tune_min_cells(0)  # time every block, however small
    trials, final = run_until_tuned(al, (0, 2999, 3000, 3255), wide_expected)
    seen = [shape(t) for t in trials]
    # stage 0: the three variants at their default tile shapes, the two-pass
    # first for a 4-code alphabet (its score pass in the max3 form: 32 rows)
    assert seen[:3] == [(ca.SCHED_PERSISTENT_I16, 32, 16), (ca.SCHED_PERSISTENT_I32, 16, 0),
                        (ca.SCHED_PERSISTENT_LPP16, 4, 0)]
    # later stages vary the winner's strips only, starting with the winner
    # itself timed again; no other configuration is measured twice
    winner = [t for t in trials[3:]]
    assert winner, "no tile-shape stage ran"
    assert len({t["variant"] for t in winner}) == 1
    assert seen[3] in seen[:3]
    assert len(set(seen[3:])) == len(seen[3:]), "a configuration was measured twice in the strip stages"
    assert shape(final) in seen and final["trials"] == len(trials)
    assert [t["trials"] for t in trials] == list(range(1, len(trials) + 1))
    # the choice sticks
    got, _, _ = al.compute_block(0, 2999, 3000, 3255, False)
    assert al.last_block_config() == final and np.array_equal(got, wide_expected)
    al.close()


def test_autotuner_with_fixed_strips_chooses_the_variant_only(wide_set, wide_expected, ednafull):
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(wide_set)
    al.set_schedule(ca.SCHED_PERSISTENT, 32)
    al.set_tune_min_cells(0)
    trials, final = run_until_tuned(al, (0, 2999, 3000, 3255), wide_expected)
    assert [shape(t) for t in trials] == [(ca.SCHED_PERSISTENT_I16, 32, 32), (ca.SCHED_PERSISTENT_I32, 32, 0),
                                          (ca.SCHED_PERSISTENT_LPP16, 4, 0)]
    assert shape(final) in [shape(t) for t in trials]
    al.close()


def test_autotuner_global_strips(wide_set, ednafull):
    al = ca.Aligner(ednafull, 16, 4, mode=ca.MODE_NW_GLOBAL)
    al.set_sequences(wide_set)
    al.set_tune_min_cells(0)
    a, b = np.meshgrid(np.arange(0, 3000), np.arange(3000, 3256), indexing="ij")
    d, _ = O.nw_pairs_mt(wide_set.residues, wide_set.offsets, a.ravel(), b.ravel(), ednafull, 16, 4,
                         threads=THREADS)
    trials, final = run_until_tuned(al, (0, 2999, 3000, 3255), d.reshape(a.shape))
    assert [shape(t) for t in trials] == [(ca.SCHED_PERSISTENT_I32, 16, 0), (ca.SCHED_PERSISTENT_I32, 8, 0)]
    assert shape(final) in [shape(t) for t in trials]
    al.close()


@pytest.mark.parametrize("reverse_strip,row_cap", [(8, 512), (16, 512), (16, 16), (16, 64)])
def test_reverse_strip_heights(wide_set, wide_expected, ednafull, reverse_strip, row_cap):
    # the two-pass schedule's reverse pass at both strip heights: unrelated
    # pairs (short bands) and related families (long bands, big rectangles);
    # small row caps send most pairs to their whole prefix instead
    say("reverse strip %d row cap %d" % (reverse_strip, row_cap))
    al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT_I16)
    al.set_reverse_strip(reverse_strip)
    al.set_reverse_row_cap(row_cap)
    al.set_sequences(wide_set)
    got, _, _ = al.compute_block(0, 2999, 3000, 3255, False)
    assert np.array_equal(got, wide_expected)
    fam = ca.synth_dna(62, 400, 4, 0.05, 0.02, 0, [300, 900])  # 4 ancestors: most pairs related
    al.set_sequences(fam)
    got, gst, _ = al.compute_block(0, 199, 200, 399, False, stats=True)
    a, b = np.meshgrid(np.arange(0, 200), np.arange(200, 400), indexing="ij")
    d, st = O.sw_pairs_mt(fam.residues, fam.offsets, a.ravel(), b.ravel(), ednafull, 16, 4, threads=THREADS,
                          want_stats=True)
    assert np.array_equal(got.ravel(), d)
    assert np.array_equal(gst.reshape(-1, 4)[:, :3], st[:, :3])
    al.close()


def test_trial_chunks_of_a_large_block(ednafull):
    """The autotuner times its first variant on a whole block, then the host
    entry point computes the next large block in trial chunks of about 4.5e6
    pairs, one further variant each, and the strip stages on whole blocks;
    every call gives the oracle's distances (checked on a seeded sample, then
    byte for byte against the first call)."""
    say("trial chunks")
    s = ca.synth_dna(62, 6400, 40, 0.10, 0.02, 0, [30, 90])
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(s)
    al.set_tune_min_cells(1e9)
    got, _, pairs = al.compute_block(0, 2399, 2400, 6399, False)
    assert pairs == 2400 * 4000
    assert al.last_block_config()["trials"] == 1
    rng = np.random.default_rng(3)
    a = rng.integers(0, 2400, 20000)
    b = rng.integers(2400, 6400, 20000)
    want, _ = O.sw_pairs_mt(s.residues, s.offsets, a, b, ednafull, 16, 4, threads=THREADS)
    assert np.array_equal(got[a, b - 2400], want)
    again, _, _ = al.compute_block(0, 2399, 2400, 6399, False)
    assert np.array_equal(again, got)
    # two variant trials in one call: a trial chunk and the rest of the block
    assert al.last_block_config()["trials"] == 3
    for _ in range(8):  # the strip stages, until the choice is made
        again, _, _ = al.compute_block(0, 2399, 2400, 6399, False)
        assert np.array_equal(again, got)
        if al.last_block_config()["tuning"] == 2:
            break
    assert al.last_block_config()["tuning"] == 2
    al.close()


def test_length_classes_and_diagonal_blocks(ednafull):
    """Two blocks whose mean lengths straddle the edge of a fixed length bin
    (2 log2(1 + m) = 11.5) share one autotuner record; a lower-mirror
    diagonal block runs a candidate only as its record's first block, later
    ones the fastest configuration timed so far; all bit-exact on seeded
    samples."""
    say("length classes and diagonal blocks")
    rng = np.random.default_rng(8)
    lens = np.concatenate([rng.integers(40, 63, 3000), rng.integers(48, 71, 3000), rng.integers(40, 71, 256)])
    seqs = [rng.choice(list(b"ACGT"), size=int(n)).astype(np.uint8).tobytes() for n in lens]
    s = ca.SequenceSet.from_strings(seqs)
    L = s.lengths.astype(np.float64)

    def lclass(r0, r1, c0, c1):
        return 2 * np.log2(1 + (L[r0:r1 + 1].sum() + L[c0:c1 + 1].sum()) / (r1 - r0 + 1 + c1 - c0 + 1))

    blocks = [(0, 2999, 6000, 6255), (3000, 5999, 6000, 6255)]
    lc = [lclass(*b) for b in blocks]
    assert round(lc[0]) != round(lc[1]) and abs(lc[0] - lc[1]) < 0.5, lc

    def check(got, r0, r1, c0, c1, diag=False):
        a = rng.integers(r0, r1 + 1, 20000)
        b = rng.integers(c0, c1 + 1, 20000)
        if diag:
            keep = a != b
            a, b = np.maximum(a, b)[keep], np.minimum(a, b)[keep]  # the lower triangle
        want, _ = O.sw_pairs_mt(s.residues, s.offsets, a, b, ednafull, 16, 4, threads=THREADS)
        assert np.array_equal(got[a - r0, b - c0], want)
        if diag:
            assert np.array_equal(got[b - r0, a - c0], want)  # and its mirror

    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(s)
    al.set_tune_min_cells(0)
    for k, blk in enumerate(blocks):
        got, _, _ = al.compute_block(*blk, False)
        cfg = al.last_block_config()
        assert cfg["tuning"] == 1 and cfg["trials"] == k + 1, cfg
        check(got, *blk)
    got, _, _ = al.compute_block(0, 1999, 0, 1999, True)
    cfg = al.last_block_config()
    assert cfg["tuning"] == 3 and cfg["trials"] == 2, cfg
    assert cfg["variant"] in (ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT_I32), cfg
    check(got, 0, 1999, 0, 1999, diag=True)
    al.close()
    # a diagonal block first: its record's first trial
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(s)
    al.set_tune_min_cells(0)
    got, _, _ = al.compute_block(0, 1999, 0, 1999, True)
    cfg = al.last_block_config()
    assert cfg["tuning"] == 1 and cfg["trials"] == 1, cfg
    check(got, 0, 1999, 0, 1999, diag=True)
    al.close()
