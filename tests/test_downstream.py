"""Downstream consumers of the distance matrix (PWC clustering, MDS).

They load the matrix file this path writes with MatrixBinaryReader
(Salsa.CoreTPL/Blas/MatrixBinaryReader.cs:191-217 for SWMPI's per-rank
partial stripes, :224-311 for a row range of the full file) and turn a stored
Int16 into a distance: PWCParallelism.getDistanceValue
(Salsa.PairwiseClusteringTPL/PWCParallelism.cs:211-218) and
SALSAParallelism.getDistanceValue (MDSasCHISQ/SALSAParallelism.cs:318-370).
These tests write matrices with the path's writers and read them back the way
those programs do; the consumers themselves are not part of this build.
"""
import os

import numpy as np
import pytest

import csharp_amd as ca


def _matrix(n, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.int16:
        m = rng.integers(0, 32768, size=(n, n)).astype(np.int16)
    else:
        m = rng.random((n, n))
    m = np.triu(m, 1)
    m = m + m.T  # symmetric, 0 diagonal, as WriteFullMatrixOnRank0 leaves it
    return m.astype(dtype)


@pytest.mark.parametrize("dtype", [np.int16, np.float64])
def test_full_file_row_ranges(tmp_path, dtype):
    n = 37
    m = _matrix(n, dtype, 1)
    path = str(tmp_path / "dist.bin")
    ca.write_matrix(path, m)
    assert os.path.getsize(path) == n * n * m.itemsize
    for start, count in ((0, n), (5, 11), (36, 1), (12, 0)):
        assert np.array_equal(ca.read_matrix_rows(path, n, start, count, dtype), m[start:start + count])
    with pytest.raises(ca.SwgError):
        ca.read_matrix_rows(path, n, 30, 8, dtype)  # past the end of the file


@pytest.mark.parametrize("world", [1, 3, 4])
def test_partial_stripes(tmp_path, world):
    n = 37
    m = _matrix(n, np.int16, 2)
    ranges = ca.range_partition(n, world)
    for i, (r0, r1) in enumerate(ranges):
        ca.write_matrix(str(tmp_path / ("dist_%d.bin" % i)), m[r0:r1 + 1])
    assert np.array_equal(ca.read_partial_stripes(str(tmp_path), "dist", ".bin", 0, world - 1, ranges, n), m)
    if world > 1:
        got = ca.read_partial_stripes(str(tmp_path), "dist", ".bin", 1, world - 1, ranges, n)
        assert np.array_equal(got, m[ranges[1][0]:])
    with pytest.raises(ca.SwgError):
        ca.read_partial_stripes(str(tmp_path), "dist", ".bin", 0, world, ranges + [(n, n)], n)


def test_consumer_distance_values():
    # PWC: value / 32767.0 over the whole Int16 range written by ComputeDistance
    for v in (0, 1, 16383, 18723, 32766, 32767):
        assert ca.distance_value(v, ca.PWC) == v / 32767.0
    # MDS: the maximum is "undefined" only with a distance cut; negative values are undefined
    assert ca.distance_value(32767, ca.MDS) == 1.0
    assert ca.distance_value(32767, ca.MDS, distance_cut=0.5) == -1.0
    assert ca.distance_value(32767, ca.MDS, distance_cut=0.5, undefined_value=3.0) == 3.0
    assert ca.distance_value(32767, ca.MDS, distance_cut=0.5, undefined_value=3.0, processing_option=2) == 9.0
    assert ca.distance_value(-2, ca.MDS) == -1.0
    assert ca.distance_value(-2, ca.MDS, undefined_value=0.25, processing_option=2) == 0.0625
    assert ca.distance_value(16383, ca.MDS) == 16383 / 32767.0
    assert ca.distance_value(16383, ca.MDS, processing_option=2) == (16383 / 32767.0) ** 2


def test_int16_distance_round_trip_through_consumer():
    """A stored Int16 distance (short)((1 - pct) * 32767) read back by PWC is
    1 - pct up to one unit of the Int16 scale (the truncation)."""
    for cnt, cols in ((0, 1), (1, 3), (3, 7), (250, 300), (999, 1000), (1000, 1000)):
        pct = np.float32(np.float64(cnt) / np.float64(cols))
        stored = int(np.float32(np.float32(1.0) - pct) * np.float32(32767.0))
        d = ca.distance_value(stored, ca.PWC)
        assert 0.0 <= (1.0 - float(pct)) - d < 1.0 / 32767.0 + 1e-12


def test_index_file(tmp_path):
    # WriteIndexFile (Salsa.SmithWatermanMPI/Program.cs:316-331): "{0}\t{1}\t{2}"
    # of index, label and residue count, one StreamWriter line each
    seqs = ca.SequenceSet.from_strings([b"ACGT", b"", b"GATTACA"], labels=["seq one", "s2", "éx"])
    p = str(tmp_path / "index.txt")
    ca.write_index_file(p, seqs)
    assert open(p, "rb").read() == "0\tseq one\t4\r\n1\ts2\t0\r\n2\téx\t7\r\n".encode("utf-8")
    ca.write_index_file(p, seqs, crlf=False)
    assert open(p, "rb").read() == "0\tseq one\t4\n1\ts2\t0\n2\téx\t7\n".encode("utf-8")
    ca.write_index_file(p, ca.SequenceSet.from_strings([]))
    assert open(p, "rb").read() == b""
