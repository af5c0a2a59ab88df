"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact comparison of Int16/Double distances and of the per-pair
statistics (score, aligned length, identity/similarity count) on seeded
inputs, over every block kind and schedule, plus edge cases.
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


def oracle_block(seqs, r0, r1, c0, c1, diag_lower, M, o, e, nucleic=True, tdouble=False):
    """Expected block (rows x cols) and stats, computing (i, j) with s1 = s_i."""
    rows, cols = r1 - r0 + 1, c1 - c0 + 1
    a, b = [], []
    for i in range(r0, r1 + 1):
        for j in range(c0, c1 + 1):
            if diag_lower and j >= i:
                continue
            a.append(i); b.append(j)
    d, st = O.sw_pairs_mt(seqs.residues, seqs.offsets, a, b, M, o, e, nucleic=nucleic, tdist_double=tdouble,
                          threads=THREADS, want_stats=True)
    out = np.zeros((rows, cols), dtype=np.float64 if tdouble else np.int16)
    stats = np.zeros((rows, cols, 4), dtype=np.int32)
    a = np.asarray(a) - r0
    b = np.asarray(b) - c0
    out[a, b] = d
    stats[a, b, 0] = st[:, 0]
    stats[a, b, 1] = st[:, 1]
    stats[a, b, 2] = st[:, 2] if nucleic else st[:, 3]
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
    assert np.array_equal(gst[..., :3], wst[..., :3]), what + ": stats differ"


@pytest.fixture(scope="module")
def dna_set():
    # related and unrelated DNA, lengths 1..260, plus edge sequences
    s = ca.synth_dna(11, 180, 12, 0.10, 0.02, 0, [1, 260])
    extra = [b"", b"A", b"ACGT" * 40, b"A" * 150, b"ACACACACAC" * 12, b"NNNNACGTNN", b"ACGUXACGT", b"T" * 300]
    allseq = [s.seq(i) for i in range(len(s))] + extra
    return ca.SequenceSet.from_strings(allseq)


@pytest.fixture(scope="module")
def protein_set():
    s = ca.synth_protein(12, 120, 2, [120, 0.5, 1, 400])
    extra = [b"", b"W", b"XXBZ*AR", b"ACDEFGHIKLMNPQRSTVWY" * 5]
    return ca.SequenceSet.from_strings([s.seq(i) for i in range(len(s))] + extra)


SCHEDULES = [(ca.SCHED_PERSISTENT, 0), (ca.SCHED_PLAIN, 16), (ca.SCHED_PERSISTENT_I32, 8),
             (ca.SCHED_PERSISTENT_I32, 16), (ca.SCHED_PERSISTENT_I32, 32), (ca.SCHED_PERSISTENT_I16, 16),
             (ca.SCHED_PERSISTENT_I16, 32), (ca.SCHED_PERSISTENT_LPP16, 0), (ca.SCHED_PERSISTENT_LPP16, 8),
             (ca.SCHED_PERSISTENT_LPS16, 0), (ca.SCHED_PERSISTENT_LPS16, 32)]
BLOCK_SCHEDULES = [ca.SCHED_PERSISTENT, ca.SCHED_PLAIN, ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16,
                   ca.SCHED_PERSISTENT_LPP16, ca.SCHED_PERSISTENT_LPS16]


@pytest.mark.parametrize("schedule,strip", SCHEDULES)
def test_offdiag_block_int16(dna_set, ednafull, schedule, strip):
    say("offdiag int16 schedule=%d strip=%d" % (schedule, strip))
    al = ca.Aligner(ednafull, 16, 4, schedule=schedule)
    al.set_schedule(schedule, strip)
    al.set_sequences(dna_set)
    n = len(dna_set)
    r0, r1, c0, c1 = 0, 89, 90, n - 1
    got, gst, pairs = al.compute_block(r0, r1, c0, c1, False, stats=True)
    want, wst = oracle_block(dna_set, r0, r1, c0, c1, False, ednafull, 16, 4)
    assert pairs == (r1 - r0 + 1) * (c1 - c0 + 1)
    assert_same(got, want, gst, wst, "offdiag")
    al.close()


@pytest.mark.parametrize("schedule", BLOCK_SCHEDULES)
def test_diagonal_block_lower_mirror(dna_set, ednafull, schedule):
    say("diag lower-mirror schedule=%d" % schedule)
    al = ca.Aligner(ednafull, 16, 4, schedule=schedule)
    al.set_sequences(dna_set)
    n = len(dna_set)
    r0, r1 = 60, n - 1
    got, gst, pairs = al.compute_block(r0, r1, r0, r1, True, stats=True)
    want, wst = oracle_block(dna_set, r0, r1, r0, r1, True, ednafull, 16, 4)
    m = r1 - r0 + 1
    assert pairs == m * (m - 1) // 2
    assert_same(got, want, gst, wst, "diag")
    assert np.all(np.diag(got) == 0)
    al.close()


@pytest.mark.parametrize("schedule", BLOCK_SCHEDULES)
def test_diagonal_block_full_swms(dna_set, ednafull, schedule):
    say("diag full schedule=%d" % schedule)
    al = ca.Aligner(ednafull, 16, 4, diag_mode=ca.DIAG_FULL, schedule=schedule)
    al.set_sequences(dna_set)
    r0, r1 = 100, 170
    got, gst, pairs = al.compute_block(r0, r1, r0, r1, True, stats=True)
    want, wst = oracle_block(dna_set, r0, r1, r0, r1, False, ednafull, 16, 4)
    assert pairs == (r1 - r0 + 1) ** 2
    assert_same(got, want, gst, wst, "diag-full")
    al.close()


def test_double_distances(dna_set, ednafull):
    say("double")
    al = ca.Aligner(ednafull, 16, 4, tdistance=ca.TDIST_DOUBLE)
    al.set_sequences(dna_set)
    n = len(dna_set)
    got, gst, _ = al.compute_block(0, 70, 71, n - 1, False, stats=True)
    want, wst = oracle_block(dna_set, 0, 70, 71, n - 1, False, ednafull, 16, 4, tdouble=True)
    assert_same(got, want, gst, wst, "double")
    al.close()


@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT_LPP16])
def test_protein_percent_similarity(protein_set, blosum62, schedule):
    say("protein schedule=%d" % schedule)
    al = ca.Aligner(blosum62, 10, 1, alignment_type=ca.ALIGN_PROTEIN, schedule=schedule)
    al.set_sequences(protein_set)
    n = len(protein_set)
    got, gst, _ = al.compute_block(0, n - 1, 0, n - 1, True, stats=True)
    want, wst = oracle_block(protein_set, 0, n - 1, 0, n - 1, True, blosum62, 10, 1, nucleic=False)
    assert_same(got, want, gst, wst, "protein")
    al.close()


def test_swapped_arguments_follow_block_order(dna_set, ednafull):
    # align(s_i, s_j) and align(s_j, s_i) may differ in identity/length:
    # the block's rows are always s1
    say("swapped")
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(dna_set)
    g1, s1, _ = al.compute_block(0, 40, 41, 90, False, stats=True)
    g2, s2, _ = al.compute_block(41, 90, 0, 40, False, stats=True)
    w1, _ = oracle_block(dna_set, 0, 40, 41, 90, False, ednafull, 16, 4)
    w2, _ = oracle_block(dna_set, 41, 90, 0, 40, False, ednafull, 16, 4)
    assert np.array_equal(g1, w1) and np.array_equal(g2, w2)
    assert np.array_equal(s1[..., 0], s2[..., 0].T)  # the score is symmetric
    al.close()


def test_full_matrix_multi_rank_layout(dna_set, ednafull):
    # every rank's checkerboard blocks + scatter mirror, P = 3 (Program.cs:618-677)
    say("full matrix P=3")
    n = 120
    sub = dna_set.subset(range(n))
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(sub)
    full, pairs = al.compute_full_matrix(world=3)
    assert pairs == n * (n - 1) // 2
    ranges = ca.range_partition(n, 3)
    owner = np.zeros(n, dtype=int)
    for r, (s, e) in enumerate(ranges):
        owner[s:e + 1] = r
    a, b = [], []
    for x in range(n):
        for y in range(x + 1, n):
            ra, rb = owner[x], owner[y]
            if ra == rb or (ra + rb) % 2 == 0:
                a.append(y); b.append(x)  # s1 = larger index
            else:
                a.append(x); b.append(y)
    d, _ = O.sw_pairs_mt(sub.residues, sub.offsets, a, b, ednafull, 16, 4, threads=THREADS)
    want = np.zeros((n, n), dtype=np.int16)
    want[a, b] = d
    want[b, a] = d
    assert np.array_equal(full, want)
    al.close()


def test_full_matrix_read_by_consumers(tmp_path, dna_set, ednafull):
    """The matrix this path produces, written as WriteFullMatrixOnRank0 and as
    SWMPI's per-rank partial stripes, loaded the way PWC and MDS load it
    (MatrixBinaryReader) and turned into their distances."""
    say("downstream consumers")
    n = 90
    sub = dna_set.subset(range(n))
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(sub)
    full, _ = al.compute_full_matrix(world=3)
    al.close()
    path = str(tmp_path / "dist.bin")
    ca.write_matrix(path, full)
    ranges = ca.range_partition(n, 3)
    for i, (r0, r1) in enumerate(ranges):
        ca.write_matrix(str(tmp_path / ("dist_%d.bin" % i)), full[r0:r1 + 1])
        assert np.array_equal(ca.read_matrix_rows(path, n, r0, r1 - r0 + 1), full[r0:r1 + 1])
    assert np.array_equal(ca.read_partial_stripes(str(tmp_path), "dist", ".bin", 0, 2, ranges, n), full)
    pwc = np.array([[ca.distance_value(v, ca.PWC) for v in row] for row in full[:20]])
    assert np.array_equal(pwc, full[:20] / 32767.0)
    assert (pwc >= 0).all() and (pwc <= 1).all() and (np.diag(pwc[:, :20]) == 0).all()
    assert np.array_equal(full, full.T)


@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT_LPP16,
                                      ca.SCHED_PERSISTENT_LPS16])
def test_long_sequences(ednafull, schedule):
    say("long sequences schedule=%d" % schedule)
    s = ca.synth_dna(13, 70, 3, 0.10, 0.02, 0, [1500, 2000])
    al = ca.Aligner(ednafull, 16, 4, schedule=schedule)
    al.set_sequences(s)
    got, gst, _ = al.compute_block(0, 5, 6, 69, False, stats=True)
    want, wst = oracle_block(s, 0, 5, 6, 69, False, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "long")
    al.close()


@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT, ca.SCHED_PERSISTENT_I32, ca.SCHED_PERSISTENT_I16,
                                      ca.SCHED_PERSISTENT_LPP16, ca.SCHED_PERSISTENT_LPS16])
def test_very_long_sequences(ednafull, schedule):
    """7,800-8,000-residue relatives (they score above the int16 range: the
    score pass flags them for the whole-matrix int32 re-run), unrelated
    sequences of the same lengths and a 5,000-residue one; then a
    16,000-residue sequence against a mutated 9,000-residue copy of its start
    and an unrelated 3,000-residue one: boundary scratch sized for the
    longest sequence, bit-exact against the oracle."""
    say("very long sequences schedule=%d" % schedule)
    rng = np.random.default_rng(41)
    fam = _mutated_family(rng, b"ACGT", 3, 7800, 8000, 0.02, 0.01)  # 0-2 related, 3-5 unrelated
    five = rng.choice(list(b"ACGT"), size=5000).astype(np.uint8).tobytes()
    big = rng.choice(list(b"ACGT"), size=16000).astype(np.uint8).tobytes()
    copy = bytearray()
    for ch in big[:9000]:
        r = rng.random()
        if r < 0.01:
            continue
        copy.append(rng.choice(list(b"ACGT")) if rng.random() < 0.05 else ch)
    other = rng.choice(list(b"ACGT"), size=3000).astype(np.uint8).tobytes()
    s = ca.SequenceSet.from_strings([fam.seq(i) for i in range(len(fam))] + [five, big, bytes(copy), other])
    al = ca.Aligner(ednafull, 16, 4, schedule=schedule)
    al.set_sequences(s)
    got, gst, _ = al.compute_block(0, 1, 2, 6, False, stats=True)
    want, wst = oracle_block(s, 0, 1, 2, 6, False, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "5,000-8,000 residues")
    assert (wst[..., 0] > 32766).any(), "a related pair must exceed the int16 range"
    got, gst, _ = al.compute_block(7, 7, 8, 9, False, stats=True)
    want, wst = oracle_block(s, 7, 7, 8, 9, False, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "16,000 residues")
    assert wst[0, 0, 0] > 32766
    al.close()


def _mutated_family(rng, alphabet, n, lo, hi, sub, indel):
    """n copies of one random ancestor with substitutions and indels, plus
    as many unrelated sequences: related pairs align over most of their
    length, so their re-run rectangles exceed the big-rectangle threshold."""
    anc = rng.choice(list(alphabet), size=hi)
    out = []
    for _ in range(n):
        s = []
        for ch in anc[: rng.integers(lo, hi + 1)]:
            r = rng.random()
            if r < indel / 2:
                continue
            if r < indel:
                s.append(rng.choice(list(alphabet)))
            s.append(rng.choice(list(alphabet)) if rng.random() < sub else ch)
        out.append(bytes(s))
    for _ in range(n):
        out.append(rng.choice(list(alphabet), size=int(rng.integers(lo, hi + 1))).astype(np.uint8).tobytes())
    return ca.SequenceSet.from_strings(out)


@pytest.mark.parametrize("kind", ["dna", "protein"])
def test_big_rectangles_one_wave_per_pair(ednafull, blosum62, kind):
    """Related pairs whose re-run rectangles exceed 65536 cells go to the
    one-wave-per-pair kernel (8 rows per lane for DNA, 4 for the protein
    alphabet); their distances and statistics must match the oracle."""
    say("big rectangles %s" % kind)
    rng = np.random.default_rng(31 if kind == "dna" else 37)
    if kind == "dna":
        s = _mutated_family(rng, b"ACGT", 24, 600, 1300, 0.10, 0.03)
        M, o, e, nucleic = ednafull, 16, 4, True
    else:
        s = _mutated_family(rng, b"ACDEFGHIKLMNPQRSTVWY", 24, 300, 700, 0.25, 0.03)
        M, o, e, nucleic = blosum62, 10, 1, False
    al = ca.Aligner(M, o, e, schedule=ca.SCHED_PERSISTENT_I16,
                    alignment_type=ca.ALIGN_NUCLEIC if nucleic else ca.ALIGN_PROTEIN)
    al.set_schedule(ca.SCHED_PERSISTENT_I16, 16)
    al.set_sequences(s)
    n = len(s)
    al.set_kernel_timing(True)
    got, gst, _ = al.compute_block(0, 23, 0, n - 1, False, stats=True)
    kt = {k[0]: k for k in al.kernel_times()}
    al.set_kernel_timing(False)
    want, wst = oracle_block(s, 0, 23, 0, n - 1, False, M, o, e, nucleic=nucleic)
    assert_same(got, want, gst, wst, "big rectangles")
    assert kt["rect64"][2] > 0 and kt["rect64"][3] >= kt["rect64"][2]
    # the graph path gives the same block
    got2, gst2, _ = al.compute_block(0, 23, 0, n - 1, False, stats=True)
    assert np.array_equal(got2, got) and np.array_equal(gst2, gst)
    al.close()


def test_repeated_blocks_reuse_graph(dna_set, ednafull):
    say("repeat")
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(dna_set)
    want, wst = oracle_block(dna_set, 0, 30, 31, 120, False, ednafull, 16, 4)
    for _ in range(3):
        got, gst, _ = al.compute_block(0, 30, 31, 120, False, stats=True)
        assert_same(got, want, gst, wst, "repeat")
        other, _, _ = al.compute_block(40, 60, 0, 39, False)
    al.close()


def test_int16_saturation_reruns_exactly():
    # match +100: identical 400-residue sequences score 40000 > 32767, so the
    # int16 score pass saturates and those pairs must be re-run in int32
    say("saturation")
    M = np.zeros((127, 127), dtype=np.float32)
    for x in b"ACGT":
        for y in b"ACGT":
            M[x, y] = 100 if x == y else -60
    s = ca.synth_dna(14, 40, 2, 0.02, 0.005, 0, [300, 420])
    seqs = ca.SequenceSet.from_strings([s.seq(i) for i in range(40)] + [s.seq(0), s.seq(0), b"ACGT" * 100])
    for sched in (ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT_I32):
        al = ca.Aligner(M, 150, 20, schedule=sched)
        al.set_sequences(seqs)
        n = len(seqs)
        got, gst, _ = al.compute_block(0, n - 1, 0, n - 1, True, stats=True)
        want, wst = oracle_block(seqs, 0, n - 1, 0, n - 1, True, M, 150, 20)
        assert wst[..., 0].max() > 32767
        assert_same(got, want, gst, wst, "saturation sched=%d" % sched)
        al.close()


def test_many_partners_multi_chunk(ednafull):
    # a row with more than 2048 partners is grouped in several chunks
    say("multi-chunk")
    s = ca.synth_dna(15, 2300, 30, 0.10, 0.02, 0, [8, 60])
    for sched in (ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT):
        al = ca.Aligner(ednafull, 16, 4, schedule=sched)
        al.set_sequences(s)
        got, gst, _ = al.compute_block(0, 7, 8, 2299, False, stats=True)
        want, wst = oracle_block(s, 0, 7, 8, 2299, False, ednafull, 16, 4)
        assert_same(got, want, gst, wst, "multi-chunk sched=%d" % sched)
        got, gst, _ = al.compute_block(2200, 2299, 2200, 2299, True, stats=True)
        want, wst = oracle_block(s, 2200, 2299, 2200, 2299, True, ednafull, 16, 4)
        assert_same(got, want, gst, wst, "multi-chunk diag sched=%d" % sched)
        al.close()


def test_autotuned_blocks_match(ednafull):
    # a block large enough for the autotuner's sample; both variants' results
    # must agree with the oracle regardless of the choice
    say("autotune")
    s = ca.synth_dna(16, 700, 20, 0.10, 0.02, 0, [60, 200])
    al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT)
    al.set_sequences(s)
    got, gst, _ = al.compute_block(0, 349, 350, 699, False, stats=True)
    want, wst = oracle_block(s, 0, 349, 350, 699, False, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "autotune offdiag")
    got, gst, _ = al.compute_block(350, 699, 350, 699, True, stats=True)
    want, wst = oracle_block(s, 350, 699, 350, 699, True, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "autotune diag")
    al.close()


@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT, ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT_LPP16,
                                      ca.SCHED_PERSISTENT_LPS16])
def test_serving_same_alone_and_in_mix(ednafull, schedule):
    # each query's distances must be identical when it runs alone and when it
    # shares batches with others, and equal to the oracle's
    say("serving alone vs mix, schedule=%d" % schedule)
    panel = ca.synth_dna(21, 300, 10, 0.10, 0.02, 0, [30, 250])
    qs = ca.synth_dna(22, 24, 10, 0.10, 0.02, 0, [1, 400])
    queries = [qs.seq(i) for i in range(len(qs))] + [b"", b"NACGU"]
    srv = ca.Server(panel, ednafull, 16, 4, max_query_len=600, max_batch=5, lanes=2, hbm_budget=64 << 20,
                    schedule=schedule)
    alone = []
    for q in queries:
        t = srv.submit(q)
        assert srv.wait(t, 60000)
        out, st = srv.result(t)
        alone.append((out.copy(), st.copy()))
        srv.release(t)
    tickets = [srv.submit(q) for q in queries]
    for t in tickets:
        assert srv.wait(t, 60000)
    for k, t in enumerate(tickets):
        out, st = srv.result(t)
        assert np.array_equal(out, alone[k][0]), "query %d differs in the mix" % k
        assert np.array_equal(st, alone[k][1])
    # oracle: s1 = the query, s2 = each panel sequence
    allseq = ca.SequenceSet.from_strings([panel.seq(i) for i in range(len(panel))] + queries)
    npan = len(panel)
    for k in range(len(queries)):
        a = np.full(npan, npan + k)
        b = np.arange(npan)
        d, st = O.sw_pairs_mt(allseq.residues, allseq.offsets, a, b, ednafull, 16, 4, threads=THREADS,
                              want_stats=True)
        out, gst = srv.result(tickets[k])
        assert np.array_equal(out, d), "query %d vs oracle" % k
        assert np.array_equal(gst[:, :3], st[:, :3])
    srv.close()


def test_streamed_block_matches(dna_set, ednafull):
    # a block larger than the stream chunk is computed chunk by chunk with the
    # copy of chunk k-1 overlapping chunk k
    say("streamed block")
    al = ca.Aligner(ednafull, 16, 4)
    al.set_sequences(dna_set)
    al.set_stream_chunk_bytes(2 * 50 * 7)   # 7 rows of 50 columns per chunk
    n = len(dna_set)
    got, _, pairs = al.compute_block(0, 99, n - 50, n - 1, False)
    want, _ = oracle_block(dna_set, 0, 99, n - 50, n - 1, False, ednafull, 16, 4)
    assert pairs == 100 * 50
    assert np.array_equal(got, want)
    al.close()


def test_swms_restart_reproduces_block_files(tmp_path, ednafull):
    # SWMS: an interrupted rank restarted from its last saved block writes the
    # same block files as an uninterrupted run (Program.cs:1164-1227, 245-262)
    say("swms restart")
    s = ca.synth_dna(23, 160, 10, 0.10, 0.02, 0, [20, 120])
    labels = ["s%03d" % i for i in range(160)]
    full = tmp_path / "full"; full.mkdir()
    part = tmp_path / "part"; part.mkdir()
    done = tmp_path / "done"; done.mkdir()
    al = ca.Aligner(ednafull, 16, 4, diag_mode=ca.DIAG_FULL)
    al.set_sequences(s)
    last_owned = {0: 3, 1: 2, 2: 3, 3: 3}   # checkerboard owners at P = 4
    for rank in range(4):
        rc, saved, pairs = al.swms_compute_rank(4, rank, str(full), str(done), labels=labels)
        assert rc == ca.SWG_OK and saved == last_owned[rank]
    # rank 0 owns blocks 0, 1, 3: stop after one block, restart from the restart file
    rc, saved, _ = al.swms_compute_rank(4, 0, str(part), None, labels=labels, stop_after=1)
    assert rc == ca.SWG_E_CAPACITY and saved == 0
    restart = tmp_path / "restart.txt"
    restart.write_text("4\n0\t%d\n" % saved)
    world, rank, last = ca.read_restart_file(str(restart), 0)
    assert (world, rank, last) == (4, 0, 0)
    rc, saved, _ = al.swms_compute_rank(world, rank, str(part), str(done), last_saved_block=last, labels=labels)
    assert rc == ca.SWG_OK
    for j in (0, 1, 3):
        a = (full / ("0-%d.bin" % j)).read_bytes()
        b = (part / ("0-%d.bin" % j)).read_bytes()
        assert a == b, "block 0-%d differs after restart" % j
    assert not (full / "0-2.bin").exists()
    assert (done / "0.txt").read_text() == "Done: Rank 0 done.\n"
    # the block payload equals ComputeBlock's output (full diagonal block, SWMS semantics)
    ranges = ca.range_partition(160, 4)
    (r0, r1), (c0, c1) = ranges[0], ranges[1]
    got, _, _ = al.compute_block(r0, r1, c0, c1, False)
    payload = (full / "0-1.bin").read_bytes()[-got.nbytes:]
    assert payload == got.tobytes()
    want, _ = oracle_block(s, r0, r1, c0, c1, False, ednafull, 16, 4)
    assert np.array_equal(got, want)
    al.close()


@pytest.mark.parametrize("schedule,names", [
    (ca.SCHED_PERSISTENT_I16, ["score16", "sort", "reverse", "sort", "rerun_track", "rerun_capture", "rect64"]),
    (ca.SCHED_PERSISTENT_LPS16, ["lps16", "sort", "reverse", "sort", "rerun_track", "rerun_capture", "rect64"]),
    (ca.SCHED_PERSISTENT_I32, ["one_pass"]), (ca.SCHED_PLAIN, ["one_pass"])])
def test_kernel_timing_mode(dna_set, ednafull, schedule, names):
    say("kernel timing schedule=%d" % schedule)
    al = ca.Aligner(ednafull, 16, 4, schedule=schedule)
    al.set_schedule(schedule, 0)
    al.set_sequences(dna_set)
    n = len(dna_set)
    ref, rst, _ = al.compute_block(0, 89, 90, n - 1, False, stats=True)
    al.set_kernel_timing(True)
    got, gst, _ = al.compute_block(0, 89, 90, n - 1, False, stats=True)
    kt = al.kernel_times()
    ms, cells, _, _ = al.last_kernel_stats()
    # the second timed block replays the same graph with its stamp kernels
    got2, _, _ = al.compute_block(0, 89, 90, n - 1, False, stats=True)
    kt2 = al.kernel_times()
    ms2, _, _, _ = al.last_kernel_stats()
    al.set_kernel_timing(False)
    assert np.array_equal(got, ref) and np.array_equal(gst, rst) and np.array_equal(got2, ref)
    assert [k[0] for k in kt] == names and [k[0] for k in kt2] == names
    assert all(k[1] > 0 for k in kt + kt2)
    # the stamps lie between the block's start and end events
    assert sum(k[1] for k in kt) <= ms * 1.02 + 0.02 and sum(k[1] for k in kt) >= 0.9 * ms - 0.1
    assert sum(k[1] for k in kt2) <= ms2 * 1.02 + 0.02 and sum(k[1] for k in kt2) >= 0.9 * ms2 - 0.1
    L = dna_set.lengths.astype(np.float64)
    block = float(L[0:90].sum() * L[90:n].sum())
    for name, _, c, run in kt:
        if name in ("score16", "one_pass", "lps16"):
            assert c == block
        if name in ("rerun_capture", "reverse"):
            assert 0 < c < block
        if name == "reverse":
            # charged with the prefix rectangles it decides; it evaluates only
            # their band (the waves' unions of the live cells)
            assert 0 < run
        else:
            assert run >= c
    al.close()


@pytest.mark.parametrize("schedule", ["two_pass", "thin"])
def test_big_rectangles_of_every_size_class(ednafull, schedule):
    """Related pairs of 300 to 3,200 residues: re-run rectangles from under
    131,072 to over 8.4e6 cells, i.e. in every size class of the big list
    (queued largest class first), 441 related pairs (not a whole number of
    64-pair tiles) beside unrelated ones; distances and statistics match the
    oracle, on the regular two-pass schedule (big above 65,536 cells) and on
    the thin-block one (big above 8,192 cells)."""
    say("big rectangles of every size class, %s" % schedule)
    rng = np.random.default_rng(41)
    lens = np.linspace(300, 3200, 21).astype(int)
    anc = rng.choice(list(b"ACGT"), size=int(lens.max()))
    fam = []
    for L in lens:
        s = []
        for ch in anc[:L]:
            r = rng.random()
            if r < 0.015:
                continue
            if r < 0.03:
                s.append(rng.choice(list(b"ACGT")))
            s.append(rng.choice(list(b"ACGT")) if rng.random() < 0.08 else ch)
        fam.append(bytes(s))
    other = [rng.choice(list(b"ACGT"), size=int(L)).astype(np.uint8).tobytes() for L in lens]
    s = ca.SequenceSet.from_strings(fam + other)
    assert min(len(x) for x in fam) ** 2 < 2 * 65536 and min(sorted(len(x) for x in fam)[-2:]) ** 2 > 128 * 65536
    if schedule == "two_pass":
        al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT_I16)
    else:
        al = ca.Aligner(ednafull, 16, 4, schedule=ca.SCHED_PERSISTENT_LPS16)
    al.set_sequences(s)
    n = len(s)
    al.set_kernel_timing(True)
    got, gst, _ = al.compute_block(0, 20, 0, n - 1, False, stats=True)
    kt = {k[0]: k for k in al.kernel_times()}
    al.set_kernel_timing(False)
    want, wst = oracle_block(s, 0, 20, 0, n - 1, False, ednafull, 16, 4)
    assert_same(got, want, gst, wst, "big rectangles of every size class")
    assert kt["rect64"][2] > 8e6
    al.close()
