"""Host-side mirror of the reference (CPU only): partitioning, checkerboard
ownership, scoring-matrix loader, FASTA parsing, file layouts, synthetic
workload determinism."""
import hashlib
import itertools
import os
import struct

import numpy as np
import pytest

import csharp_amd as ca


def py_range_partition(start, length, n):
    # RangePartitioner.Partition, Salsa.CoreTPL/RangePartitioner.cs:35-52
    chunk, rem = divmod(length, n)
    out = []
    for _ in range(n):
        if rem > 0:
            out.append((start, start + chunk)); start += chunk + 1; rem -= 1
        else:
            out.append((start, start + chunk - 1)); start += chunk
    return out


@pytest.mark.parametrize("length,n", [(10, 3), (1000, 7), (5, 5), (3, 5), (100000, 4), (0, 2)])
def test_range_partition(length, n):
    got = ca.range_partition(length, n)
    assert got == py_range_partition(0, length, n)
    # ranges tile [0, length) in order; sizes differ by at most one
    pos = 0
    for s, e in got:
        assert s == pos
        pos = e + 1
    assert pos == length
    sizes = [e - s + 1 for s, e in got]
    assert max(sizes) - min(sizes) <= 1


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5, 8])
def test_checkerboard_covers_every_pair_once(P):
    # every unordered block pair {r, j} is computed by exactly one of its two owners
    for r, j in itertools.product(range(P), repeat=2):
        if r == j:
            assert ca.block_is_computed(r, j)
        else:
            assert ca.block_is_computed(r, j) + ca.block_is_computed(j, r) == 1


def test_scoring_matrices(ednafull, blosum62):
    A, T, G, C, U = (ord(c) for c in "ATGCU")
    # EDNAFULL row A (Matrices/EDNAFULL:10) and the T/U entries
    assert [ednafull[A, x] for x in (A, T, G, C)] == [5, -4, -4, -4]
    assert ednafull[T, U] == 5 and ednafull[U, T] == 5
    assert ednafull[ord("N"), ord("N")] == -1
    W, star = ord("W"), ord("*")
    assert blosum62[W, W] == 11 and blosum62[star, star] == 1 and blosum62[ord("A"), ord("R")] == -1
    for M in (ednafull, blosum62):
        assert np.array_equal(M, M.T)
        assert np.all(M == np.round(M))
    # unlisted symbols score 0
    assert ednafull[ord("X"), ord("A")] == 0 and blosum62[ord("J"), ord("J")] == 0


def test_fasta_parser():
    text = ">seq one \r\nacgt\n  ggCC  \n>two\nNNNN\n\n>empty\n>last\nA C\n"
    s = ca.parse_fasta(text)
    assert s.labels == ["seq one", "two", "empty", "last"]
    assert [s.seq(i) for i in range(len(s))] == [b"ACGTGGCC", b"NNNN", b"", b"A C"]


def test_swms_block_file_layout(tmp_path):
    data = np.arange(6, dtype=np.int16).reshape(2, 3)
    p = str(tmp_path / "0-1.bin")
    ca.write_swms_block(p, "proj", 0, "sys", "in.fa", 0, 1, (0, 1), (2, 4), ["r0", "r1", "c0", "c4"], data)
    b = open(p, "rb").read()
    # Salsa.SmithWatermanMS/Program.cs:1116-1160, then values row-major
    want = struct.pack("<H4sH H3s H5s HH IIII", 4, b"proj", 0, 3, b"sys", 5, b"in.fa", 0, 1, 0, 1, 2, 4)
    for nm in (b"r0", b"r1", b"c0", b"c4"):
        want += struct.pack("<H", len(nm)) + nm
    want += data.tobytes()
    assert b == want


def test_synthetic_workloads_are_deterministic():
    a = ca.synth_dna(1, 200, 20, 0.10, 0.02, 0, [200, 600])
    b = ca.synth_dna(1, 200, 20, 0.10, 0.02, 0, [200, 600])
    assert np.array_equal(a.residues, b.residues) and np.array_equal(a.offsets, b.offsets)
    assert set(np.unique(a.residues).tolist()) <= set(b"ACGT")
    L = a.lengths
    assert 150 < L.mean() < 500
    p = ca.synth_protein(2, 300, 2, [300, 0.5, 50, 2000])
    assert set(np.unique(p.residues).tolist()) <= set(b"ARNDCQEGHILKMFPSTWYV")
    assert 50 <= p.lengths.min() and p.lengths.max() <= 2000
    bm = ca.synth_dna(3, 2000, 200, 0.15, 0.03, 1, [0.6, 450, 50, 1450, 100, 100, 2000])
    Lb = bm.lengths
    assert 0.5 < (Lb < 900).mean() < 0.7
    h = hashlib.sha256(a.residues.tobytes()).hexdigest()
    assert h == hashlib.sha256(ca.synth_dna(1, 200, 20, 0.10, 0.02, 0, [200, 600]).residues.tobytes()).hexdigest()
