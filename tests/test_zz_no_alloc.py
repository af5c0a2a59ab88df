"""The allocation-free entry point (swg_compute_block_device_noalloc) against
the allocating one (swg_compute_block_device) through the C ABI: bit-exact
distances and statistics on two calls in a row, the second on a smaller
block, for the distance kinds and block kinds it serves; a sample checked
against the CPU oracle; its capacity limit and unsupported mode refused with
a clear error."""
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
    s = ca.synth_dna(31, 260, 12, 0.10, 0.02, 0, [1, 300])
    extra = [b"", b"A", b"ACGT" * 50, b"T" * 220]
    return ca.SequenceSet.from_strings([s.seq(i) for i in range(len(s))] + extra)


@pytest.fixture(scope="module")
def protein_set():
    s = ca.synth_protein(32, 150, 2, [150, 0.5, 1, 400])
    return ca.SequenceSet.from_strings([s.seq(i) for i in range(len(s))] + [b"", b"W"])


def run_pair(al, blk, diag, tdt, ld=None):
    """(noalloc distances, stats), (allocating distances, stats) of one block
    into device buffers pre-filled with different garbage."""
    import torch
    r0, r1, c0, c1 = blk
    rows, cols = r1 - r0 + 1, c1 - c0 + 1
    ld = ld or cols
    res = []
    for k, fn in enumerate((al.compute_block_device_noalloc, al.compute_block_device)):
        out = torch.full((rows, ld), 111 + k, dtype=tdt, device="cuda")
        st = torch.full((rows, ld, 4), -5 - k, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        pairs = fn(r0, r1, c0, c1, diag, out.data_ptr(), ld, st.data_ptr())
        al.synchronize()
        res.append((out.cpu().numpy()[:, :cols], st.cpu().numpy()[:, :cols], pairs))
    return res


def check_same(res, what):
    (d0, s0, p0), (d1, s1, p1) = res
    assert p0 == p1, what + ": pair counts differ"
    same = (d0 == d1) | (np.isnan(d0) & np.isnan(d1)) if d0.dtype == np.float64 else d0 == d1
    assert same.all(), "%s: %d distances differ" % (what, (~same).sum())
    assert np.array_equal(s0, s1), what + ": statistics differ"


@pytest.mark.parametrize("dist,program,tdist", [
    (ca.DIST_PERCENT_IDENTITY, ca.PROGRAM_SWMPI, ca.TDIST_INT16),
    (ca.DIST_PERCENT_IDENTITY, ca.PROGRAM_SWMPI, ca.TDIST_DOUBLE),
    (ca.DIST_KIMURA2, ca.PROGRAM_SWMS, ca.TDIST_INT16),
    (ca.DIST_JUKES_CANTOR, ca.PROGRAM_SWMS, ca.TDIST_INT16),
    (ca.DIST_MINMAX_NORM, ca.PROGRAM_SWMS, ca.TDIST_INT16),
])
def test_noalloc_matches_device_entry(dna_set, ednafull, dist, program, tdist):
    import torch
    tdt = torch.float64 if tdist == ca.TDIST_DOUBLE else torch.int16
    say("noalloc dna dist=%d program=%d tdist=%d" % (dist, program, tdist))
    al = ca.Aligner(ednafull, 16, 4, distance_fn=dist, tdistance=tdist, program=program)
    al.set_sequences(dna_set)
    n = len(dna_set)
    # a large off-diagonal block, then a smaller lower-mirror diagonal block,
    # then a single-row diagonal block (no tasks: only its (i, i))
    big = run_pair(al, (0, 129, 130, n - 1), False, tdt, ld=n - 130 + 7)
    check_same(big, "off-diagonal block")
    small = run_pair(al, (200, 263, 200, 263), True, tdt)
    check_same(small, "diagonal block")
    assert (np.diagonal(small[0][0]) == 0).all()
    one = run_pair(al, (5, 5, 5, 5), True, tdt)
    check_same(one, "one-row diagonal block")
    # a sample of the big block against the oracle
    rng = np.random.default_rng(3)
    a = rng.integers(0, 130, size=300)
    b = rng.integers(130, n, size=300)
    want, _ = O.sw_pairs_mt(dna_set.residues, dna_set.offsets, a, b, ednafull, 16, 4, dist_fn=dist,
                            program=program, tdist_double=tdist == ca.TDIST_DOUBLE, threads=THREADS)
    got = big[0][0][a, b - 130]
    if tdist == ca.TDIST_DOUBLE:
        assert np.allclose(got, want, rtol=4 * np.finfo(np.float64).eps, atol=0, equal_nan=True)
    else:
        assert np.array_equal(got, want)
    al.close()


def test_noalloc_protein_and_strip(protein_set, blosum62):
    import torch
    say("noalloc protein")
    al = ca.Aligner(blosum62, 10, 1, alignment_type=ca.ALIGN_PROTEIN)
    al.set_sequences(protein_set)
    n = len(protein_set)
    for strip in (0, 8):  # 4 or 8 rows per lane
        al.set_schedule(ca.SCHED_PERSISTENT, strip)
        check_same(run_pair(al, (0, 70, 71, n - 1), False, torch.int16), "protein block strip %d" % strip)
        check_same(run_pair(al, (80, 120, 80, 120), True, torch.int16), "protein diagonal strip %d" % strip)
    al.close()


def test_noalloc_limits(ednafull):
    import torch
    say("noalloc limits")
    # Kimura2 keeps two count words per pair: 3000 x 3000 pairs need 72 MB of
    # count words alone, beyond the 64 MiB scratch
    seqs = ca.synth_dna(33, 3000, 10, 0.1, 0.02, 0, [20, 40])
    al = ca.Aligner(ednafull, 16, 4, distance_fn=ca.DIST_KIMURA2)
    al.set_sequences(seqs)
    out = torch.zeros((3000, 3000), dtype=torch.int16, device="cuda")
    with pytest.raises(ca.SwgError) as ei:
        al.compute_block_device_noalloc(0, 2999, 0, 2999, False, out.data_ptr(), 3000)
    assert ei.value.code == ca.SWG_E_CAPACITY and "MiB" in str(ei.value)
    # the context still works, on the same entry point, for a block that fits
    p = al.compute_block_device_noalloc(0, 99, 100, 299, False, out.data_ptr(), 3000)
    al.synchronize()
    assert p == 100 * 200
    ref = torch.zeros_like(out)
    al.compute_block_device(0, 99, 100, 299, False, ref.data_ptr(), 3000)
    al.synchronize()
    assert torch.equal(out[:100, :200], ref[:100, :200])
    al.close()
    # the global alignment has no allocation-free kernel
    g = ca.Aligner(ednafull, 16, 4, mode=ca.MODE_NW_GLOBAL)
    g.set_sequences(seqs)
    with pytest.raises(ca.SwgError) as ei:
        g.compute_block_device_noalloc(0, 9, 10, 19, False, out.data_ptr(), 3000)
    assert ei.value.code == ca.SWG_E_UNSUPPORTED
    g.close()
