"""GPU parity (through the C ABI) across alphabet sizes: the int16 score pass
uses a joint profile of both packed pairs' scores up to 8 codes (pad code
included) and a per-code profile beyond, here up to the 16 IUPAC nucleotide
symbols of EDNAFULL; every size must give the oracle's distances and
statistics bit for bit."""

import numpy as np
import pytest

import csharp_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = O.cpu_threads()
IUPAC = b"ACGTRYKMSWBDHVNU"


def family(rng, alphabet, n, lo, hi, sub):
    anc = rng.choice(list(alphabet), size=hi)
    out = []
    for _ in range(n):
        L = int(rng.integers(lo, hi + 1))
        s = bytes(rng.choice(list(alphabet)) if rng.random() < sub else c for c in anc[:L])
        out.append(s)
    return out


@pytest.mark.parametrize("ncodes", [4, 8, 12, 16])
@pytest.mark.parametrize("schedule", [ca.SCHED_PERSISTENT_I16, ca.SCHED_PERSISTENT])
def test_alphabet_sizes(ednafull, ncodes, schedule):
    print("alphabet %d schedule %d" % (ncodes, schedule), flush=True)
    rng = np.random.default_rng(100 + ncodes)
    alphabet = IUPAC[:ncodes]
    seqs = family(rng, alphabet, 40, 60, 400, 0.15) + \
        [rng.choice(list(alphabet), size=int(rng.integers(1, 300))).astype(np.uint8).tobytes() for _ in range(40)]
    ss = ca.SequenceSet.from_strings(seqs)
    al = ca.Aligner(ednafull, 16, 4, schedule=schedule)
    al.set_sequences(ss)
    n = len(ss)
    got, gst, _ = al.compute_block(0, 39, 40, n - 1, False, stats=True)
    a, b = np.meshgrid(np.arange(0, 40), np.arange(40, n), indexing="ij")
    d, st = O.sw_pairs_mt(ss.residues, ss.offsets, a.ravel(), b.ravel(), ednafull, 16, 4, threads=THREADS,
                          want_stats=True)
    assert np.array_equal(got.ravel(), d)
    assert np.array_equal(gst.reshape(-1, 4)[:, :3], st[:, :3])
    # and the related pairs among themselves (lower-mirror diagonal block)
    got, gst, _ = al.compute_block(0, 39, 0, 39, True, stats=True)
    a, b = np.tril_indices(40, -1)
    d, st = O.sw_pairs_mt(ss.residues, ss.offsets, a, b, ednafull, 16, 4, threads=THREADS, want_stats=True)
    assert np.array_equal(got[a, b], d) and np.array_equal(got[b, a], d)
    assert np.array_equal(gst[a, b][:, :3], st[:, :3])
    al.close()
