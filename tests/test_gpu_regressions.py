"""GPU regression tests for defects found by review, each checked against the
CPU oracle through the C ABI.

* A block's captured graph is reused for a later block of another shape
  whose buffers did not need to grow: the distance kernel of the word-mode
  distances (Jukes-Cantor, Kimura2, MinMaxNormScore) must cover the later
  block's rows and columns.
* The device entry point on a lower-mirror diagonal block writes 0 at (i, i)
  (distance and statistics), whatever the device buffer held before.
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


@pytest.fixture(scope="module")
def dna_set():
    return ca.synth_dna(21, 320, 16, 0.10, 0.02, 0, [40, 200])


def expected(seqs, r0, r1, c0, c1, M, o, e, glob, **kw):
    a, b = np.meshgrid(np.arange(r0, r1 + 1), np.arange(c0, c1 + 1), indexing="ij")
    fn = O.nw_pairs_mt if glob else O.sw_pairs_mt
    d, st = fn(seqs.residues, seqs.offsets, a.ravel(), b.ravel(), M, o, e, threads=THREADS, want_stats=True, **kw)
    return d.reshape(a.shape), st.reshape(a.shape + (4,))


# (mode, distance, program, schedule): the word-mode distances on the local
# alignment (SWMS helpers: no exception) under both statistics schedules, and
# on the global alignment (its only kernel)
CASES = [
    (ca.MODE_SW_LOCAL, ca.DIST_KIMURA2, ca.PROGRAM_SWMS, ca.SCHED_PERSISTENT_I16),
    (ca.MODE_SW_LOCAL, ca.DIST_KIMURA2, ca.PROGRAM_SWMS, ca.SCHED_PERSISTENT_I32),
    (ca.MODE_SW_LOCAL, ca.DIST_JUKES_CANTOR, ca.PROGRAM_SWMS, ca.SCHED_PERSISTENT_I16),
    (ca.MODE_SW_LOCAL, ca.DIST_MINMAX_NORM, ca.PROGRAM_SWMS, ca.SCHED_PERSISTENT_I32),
    (ca.MODE_NW_GLOBAL, ca.DIST_JUKES_CANTOR, ca.PROGRAM_SWMPI, ca.SCHED_PERSISTENT_I32),
    (ca.MODE_NW_GLOBAL, ca.DIST_MINMAX_NORM, ca.PROGRAM_SWMPI, ca.SCHED_PERSISTENT_I32),
]


@pytest.mark.parametrize("mode,dist,program,schedule", CASES)
def test_word_distance_block_shapes_share_a_graph(dna_set, ednafull, mode, dist, program, schedule):
    # 10 x 256 and 40 x 64 blocks have the same pair count, so every buffer of
    # the context keeps its capacity and the graph captured for the first
    # shape on a buffer set is replayed for the second
    glob = mode == ca.MODE_NW_GLOBAL
    say("word distance shapes mode=%d dist=%d program=%d schedule=%d" % (mode, dist, program, schedule))
    al = ca.Aligner(ednafull, 16, 4, distance_fn=dist, schedule=schedule, mode=mode, program=program)
    al.set_sequences(dna_set)
    wide = (0, 9, 40, 295)
    tall = (0, 39, 40, 103)
    kw = dict(dist_fn=dist) if glob else dict(dist_fn=dist, program=program)
    for k, blk in enumerate((wide, wide, tall, tall)):
        got, gst, pairs = al.compute_block(*blk, False, stats=True)
        want, wst = expected(dna_set, *blk, ednafull, 16, 4, glob, **kw)
        assert pairs == want.size
        assert np.array_equal(gst[..., 0], wst[..., 0]), "block %d %s: scores differ" % (k, blk)
        bad = np.argwhere(got != want)
        assert len(bad) == 0, "block %d %s: %d distances differ, first %s got %s want %s" % (
            k, blk, len(bad), bad[:3].tolist(), [got[tuple(i)] for i in bad[:3]], [want[tuple(i)] for i in bad[:3]])
    al.close()


@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT, ca.SCHED_PLAIN, ca.SCHED_PERSISTENT_I32,
                                      ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT_LPP16])
@pytest.mark.parametrize("tdist", [ca.TDIST_INT16, ca.TDIST_DOUBLE])
def test_device_mirror_block_writes_zero_diagonal(dna_set, ednafull, schedule, tdist):
    import torch
    say("device mirror diagonal schedule=%d tdist=%d" % (schedule, tdist))
    al = ca.Aligner(ednafull, 16, 4, tdistance=tdist, schedule=schedule)
    al.set_sequences(dna_set)
    r0, r1 = 100, 219
    n = r1 - r0 + 1
    tdt = torch.float64 if tdist == ca.TDIST_DOUBLE else torch.int16
    dev = torch.full((n, n), 1234, dtype=tdt, device="cuda")
    dst = torch.full((n, n, 4), -77, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    pairs = al.compute_block_device(r0, r1, r0, r1, True, dev.data_ptr(), n, dst.data_ptr())
    al.synchronize()
    assert pairs == n * (n - 1) // 2
    got = dev.cpu().numpy()
    gst = dst.cpu().numpy()
    assert (np.diagonal(got) == 0).all(), "stale diagonal distances"
    assert (gst[np.arange(n), np.arange(n)] == 0).all(), "stale diagonal statistics"
    # the rest equals the host entry point's (zero-filled staging) block
    want, wst, _ = al.compute_block(r0, r1, r0, r1, True, stats=True)
    assert np.array_equal(got, want) and np.array_equal(gst, wst)
    al.close()
