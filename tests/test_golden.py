"""The oracle reproduces the committed golden fixtures (tests/golden/pairs_v1.npz,
written by tests/golden/make_golden.py) bit for bit: a change to the oracle's
alignment, statistics or distance epilogues shows here first.  The fixtures
are the oracle's own outputs (the reference ships no golden vectors); what
pins the oracle to the reference is test_oracle.py, test_kimura_oracle.py and
test_program_helpers_oracle.py."""
import os

import numpy as np
import pytest

import csharp_amd as ca
from tests.golden import make_golden as MG

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pairs_v1.npz")


def load():
    z = np.load(GOLDEN)  # plain arrays, no pickles
    names = bytes(z["names"]).decode().split(",")
    sets = {"dna": ca.SequenceSet(z["dna_residues"], z["dna_offsets"]),
            "protein": ca.SequenceSet(z["protein_residues"], z["protein_offsets"])}
    return z, names, sets


def cases():
    return [c[0] for c in MG.CONFIGS]


@pytest.mark.parametrize("name", cases())
def test_oracle_reproduces_golden(name):
    z, names, sets = load()
    assert name in names
    o, e, glob, program, nucleic, dist_fn, double = z[name + "_config"].tolist()
    seqs = sets[bytes(z[name + "_alphabet"]).decode()]
    M = ca.load_scoring_matrix(bytes(z[name + "_matrix"]).decode())
    d, st = MG.compute(seqs, M, o, e, glob, program, nucleic, dist_fn, double)
    want = z[name + "_dist"]
    if double:
        assert np.array_equal(d.view(np.int64), want.view(np.int64))
    else:
        assert np.array_equal(d, want)
    assert np.array_equal(st, z[name + "_stats"])
