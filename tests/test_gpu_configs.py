"""GPU parity on the five benchmark workloads themselves (csharp_amd.workloads,
seeded synthetic data at their full sizes), each run through the HIP path the
way the reference's programs run it and compared with the CPU oracle:

C1  1,000 DNA, Double: the whole matrix through the checkerboard ranks and
    the scatter mirror (Salsa.SmithWatermanMPI/Program.cs:203-314,618-677);
    every one of the 499,500 pairs checked.
C2  10,000 proteins, BLOSUM62 10/1: the whole Int16 matrix with the local
    alignment (PercentSimilarity) and with the global one
    (Salsa.NeedlemanWunschTPL/Program.cs:206-300); 20,000 seeded pairs each.
C3  50,000 bimodal DNA at P = 16: one off-diagonal block and the diagonal
    block (lower mirror, Salsa.SmithWatermanMPI/Program.cs:209-231) of rank 0;
    30,000 seeded pairs each.
C4  100,000 DNA, rank 0 of P = 4: its 25,000 x 25,000 diagonal block and the
    off-diagonal block (0, 1) through the host entry point (the off-diagonal
    one streamed in row chunks to pinned memory), written as SWMS block files
    (Salsa.SmithWatermanMS/Program.cs:1094-1160) and read back; 4,000 seeded
    pairs per block checked.
C5  the 20,000-sequence serving panel: five queries (one of 1,990 residues)
    through the server; one query checked against the whole panel, the
    others on 2,000 seeded panel sequences each.
"""
import os
import sys

import numpy as np
import pytest

import csharp_amd as ca
from csharp_amd import workloads as WL
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = O.cpu_threads()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle(seqs, a, b, name, glob=False):
    cfg = WL.CONFIGS[name]
    M, o, e = WL.scoring(name)
    fn = O.nw_pairs_mt if glob else O.sw_pairs_mt
    d, _ = fn(seqs.residues, seqs.offsets, a, b, M, o, e, nucleic=cfg["alignment_type"] == ca.ALIGN_NUCLEIC,
              tdist_double=cfg["tdistance"] == ca.TDIST_DOUBLE, threads=THREADS)
    return d


def aligner(name, glob=False):
    cfg = WL.CONFIGS[name]
    M, o, e = WL.scoring(name)
    return ca.Aligner(M, o, e, alignment_type=cfg["alignment_type"], tdistance=cfg["tdistance"],
                      mode=ca.MODE_NW_GLOBAL if glob else ca.MODE_SW_LOCAL)


def mismatches(got, want):
    if got.dtype == np.float64:
        return int((~((got == want) | (np.isnan(got) & np.isnan(want)))).sum())
    return int((got != want).sum())


def lower_pairs(rng, lo, hi, k):
    """k seeded pairs (a, b) with lo <= b < a <= hi (s1 = the larger index)."""
    a = rng.integers(lo + 1, hi + 1, size=k)
    b = lo + (rng.random(k) * (a - lo)).astype(np.int64)
    return a, b


def test_c1_full_double_matrix():
    seqs = WL.sequences("c1")
    n = len(seqs)
    al = aligner("c1")
    al.set_sequences(seqs)
    say("c1: full matrix, 4 ranks")
    full, pairs = al.compute_full_matrix(world=4)
    al.close()
    assert pairs == n * (n - 1) // 2 == 499500
    assert full.dtype == np.float64 and (np.diagonal(full) == 0).all()
    assert np.array_equal(full, full.T, equal_nan=True), "scatter mirror"
    # s1 of each pair is the row sequence of the block that computed it
    # (Program.cs:628-639): rank max(ra, rb) for a diagonal block or an even
    # rank sum (s1 = the larger index), rank min(ra, rb) for an odd one
    owner = np.zeros(n, dtype=np.int64)
    for r, (s, e) in enumerate(ca.range_partition(n, 4)):
        owner[s:e + 1] = r
    x, y = np.triu_indices(n, 1)
    ra, rb = owner[x], owner[y]
    larger_first = (ra == rb) | ((ra + rb) % 2 == 0)
    a = np.where(larger_first, y, x)
    b = np.where(larger_first, x, y)
    assert (~larger_first).any() and larger_first.any()
    say("c1: oracle on all %d pairs, %d threads" % (len(a), THREADS))
    want = oracle(seqs, a, b, "c1")
    assert mismatches(full[a, b], want) == 0


@pytest.mark.parametrize("glob", [False, True], ids=["sw", "nw"])
def test_c2_protein_matrix(glob):
    seqs = WL.sequences("c2")
    n = len(seqs)
    al = aligner("c2", glob)
    al.set_sequences(seqs)
    say("c2 %s: full 10k x 10k matrix" % ("nw" if glob else "sw"))
    full, pairs = al.compute_full_matrix(world=1)
    al.close()
    assert pairs == n * (n - 1) // 2
    assert (np.diagonal(full) == 0).all() and np.array_equal(full, full.T)
    rng = np.random.default_rng(202 + glob)
    a, b = lower_pairs(rng, 0, n - 1, 20000)
    want = oracle(seqs, a, b, "c2", glob)
    assert mismatches(full[a, b], want) == 0


def test_c3_blocks_of_rank0():
    seqs = WL.sequences("c3")
    n = len(seqs)
    world = 16
    ranges = ca.range_partition(n, world)
    j = next(j for j in range(1, world) if ca.block_is_computed(0, j))
    al = aligner("c3")
    al.set_sequences(seqs)
    rng = np.random.default_rng(303)
    (r0, r1), (c0, c1) = ranges[0], ranges[j]
    say("c3: off-diagonal block (0, %d)" % j)
    blk, _, pairs = al.compute_block(r0, r1, c0, c1, False)
    assert pairs == (r1 - r0 + 1) * (c1 - c0 + 1)
    a = rng.integers(r0, r1 + 1, size=30000)
    b = rng.integers(c0, c1 + 1, size=30000)
    assert mismatches(blk[a - r0, b - c0], oracle(seqs, a, b, "c3")) == 0
    say("c3: diagonal block (0, 0)")
    blk, _, pairs = al.compute_block(r0, r1, r0, r1, True)
    al.close()
    m = r1 - r0 + 1
    assert pairs == m * (m - 1) // 2
    assert (np.diagonal(blk) == 0).all() and np.array_equal(blk, blk.T)
    a, b = lower_pairs(rng, r0, r1, 30000)
    assert mismatches(blk[a - r0, b - r0], oracle(seqs, a, b, "c3")) == 0


def test_c4_stripe_blocks_as_swms_files(tmp_path):
    seqs = WL.sequences("c4")
    n = len(seqs)
    world, rank = 4, 0
    ranges = ca.range_partition(n, world)
    (r0, r1) = ranges[rank]
    labels = ["seq%d" % i for i in range(n)]
    al = aligner("c4")
    al.set_sequences(seqs)
    rng = np.random.default_rng(404)
    for j in (0, 1):
        assert ca.block_is_computed(rank, j)
        c0, c1 = ranges[j]
        say("c4: block (%d, %d)" % (rank, j))
        blk, _, pairs = al.compute_block(r0, r1, c0, c1, j == rank)
        rows, cols = r1 - r0 + 1, c1 - c0 + 1
        assert pairs == (rows * (rows - 1) // 2 if j == rank else rows * cols)
        path = str(tmp_path / ("%d-%d.bin" % (rank, j)))
        ca.write_swms_block(path, "c4", rank, "mi355x", "c4.fasta", rank, j, (r0, r1), (c0, c1),
                            [labels[r0], labels[r1], labels[c0], labels[c1]], blk)
        del blk
        # the block file's payload: the block's Int16 values row-major after the header
        payload = np.memmap(path, dtype=np.int16, mode="r", offset=os.path.getsize(path) - rows * cols * 2,
                            shape=(rows, cols))
        if j == rank:
            a, b = lower_pairs(rng, r0, r1, 4000)
            assert (payload[np.arange(rows), np.arange(rows)] == 0).all()
            assert np.array_equal(payload[a - r0, b - c0], payload[b - r0, a - c0]), "mirror"
        else:
            a = rng.integers(r0, r1 + 1, size=4000)
            b = rng.integers(c0, c1 + 1, size=4000)
        assert mismatches(np.asarray(payload[a - r0, b - c0]), oracle(seqs, a, b, "c4")) == 0
        del payload
        os.remove(path)
    al.close()


def test_c5_queries_against_the_panel():
    panel = WL.sequences("c5_panel")
    qset = WL.sequences("c5_queries")
    M, o, e = WL.scoring("c5_panel")
    L = qset.lengths
    qidx = [int(np.argmax(L)), int(np.argmin(np.abs(L - 300))), 0, 1, 2]
    assert L[qidx[0]] >= 1990
    qs = [qset.seq(i) for i in qidx]
    srv = ca.Server(panel, M, o, e, max_query_len=int(L.max()), max_batch=8, lanes=2, hbm_budget=4 << 30,
                    stats=False)
    say("c5: %d queries against %d panel sequences" % (len(qs), len(panel)))
    tickets = [srv.submit(q) for q in qs]
    got = []
    for t in tickets:
        assert srv.wait(t, 120000)
        got.append(srv.result(t)[0].copy())
        srv.release(t)
    srv.close()
    # the queries and the panel as one sequence set: query k is s1, panel j is s2
    both = ca.SequenceSet.from_strings(qs + [panel.seq(j) for j in range(len(panel))])
    rng = np.random.default_rng(505)
    for k in range(len(qs)):
        cols = np.arange(len(panel)) if k == 1 else np.sort(rng.choice(len(panel), 2000, replace=False))
        want = oracle(both, np.full(len(cols), k), len(qs) + cols, "c5_panel")
        assert mismatches(got[k][cols], want) == 0, "query %d (%d residues)" % (k, len(qs[k]))
