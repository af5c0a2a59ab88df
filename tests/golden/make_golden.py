"""Writes tests/golden/pairs_v1.npz: small all-pairs fixtures computed by the
CPU oracle (oracle/swg_oracle.c), one per distance configuration of the path.

The reference (C#, .NET 4.0) ships no tests or golden vectors and cannot run
here, so these fixtures are the oracle's outputs on fixed inputs: they pin the
oracle against later changes (tests/test_golden.py) and give the GPU parity
tests a committed expectation (tests/test_gpu_golden.py).  The input
sequences are stored in the file, so regenerating is never needed to use it.

    python3 tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import csharp_amd as ca  # noqa: E402
from oracle import oracle as O  # noqa: E402

# (name, alphabet, matrix, gap open, gap ext, global, program, nucleic, dist_fn, double)
CONFIGS = [
    ("swmpi_pid_i16", "dna", "EDNAFULL", 16, 4, 0, 0, 1, 0, 0),
    ("swmpi_pid_f64", "dna", "EDNAFULL", 16, 4, 0, 0, 1, 0, 1),
    ("swmpi_kimura2_i16", "dna", "EDNAFULL", 16, 4, 0, 0, 1, 1, 0),
    ("swmpi_kimura2_f64", "dna", "EDNAFULL", 16, 4, 0, 0, 1, 1, 1),
    ("swmpi_jukes_cantor_i16", "dna", "EDNAFULL", 16, 4, 0, 0, 1, 2, 0),
    ("swmpi_jukes_cantor_f64", "dna", "EDNAFULL", 16, 4, 0, 0, 1, 2, 1),
    ("swmpi_similarity_i16", "protein", "BLOSUM62", 10, 1, 0, 0, 0, 0, 0),
    ("swms_similarity_i16", "protein", "BLOSUM62", 10, 1, 0, 1, 0, 0, 0),
    ("swms_minmax_i16", "dna", "EDNAFULL", 16, 4, 0, 1, 1, 3, 0),
    ("swms_kimura2_i16", "dna", "EDNAFULL", 16, 4, 0, 1, 1, 1, 0),
    ("nw_pid_i16", "dna", "EDNAFULL", 16, 4, 1, 0, 1, 0, 0),
    ("nw_similarity_i16", "protein", "BLOSUM62", 10, 1, 1, 0, 0, 0, 0),
    ("nw_minmax_i16", "dna", "EDNAFULL", 16, 4, 1, 0, 1, 3, 0),
    ("nw_jukes_cantor_f64", "dna", "EDNAFULL", 16, 4, 1, 0, 1, 2, 1),
]


def sequences():
    dna = ca.synth_dna(101, 28, 4, 0.12, 0.03, 0, [1, 180])
    dna_extra = [b"", b"A", b"ACGT" * 12, b"T" * 60, b"AGAGAGAGAG" * 4, b"NNACGTNN", b"ACGTXACGTR"]
    prot = ca.synth_protein(102, 18, 2, [120, 0.5, 1, 200])
    prot_extra = [b"", b"W", b"WWXWW", b"XXBZ*AR", b"ACDEFGHIKLMNPQRSTVWYX" * 2]
    return (ca.SequenceSet.from_strings([dna.seq(i) for i in range(len(dna))] + dna_extra),
            ca.SequenceSet.from_strings([prot.seq(i) for i in range(len(prot))] + prot_extra))


def compute(seqs, M, o, e, glob, program, nucleic, dist_fn, double):
    n = len(seqs)
    a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    d, st = O.sw_pairs_mt(seqs.residues, seqs.offsets, a.ravel(), b.ravel(), M, o, e, nucleic=bool(nucleic),
                          dist_fn=dist_fn, tdist_double=bool(double), threads=8, want_stats=True,
                          program=program, global_=bool(glob))
    return d.reshape(n, n), st.reshape(n, n, 4)


def main():
    dna, prot = sequences()
    out = {"dna_residues": dna.residues, "dna_offsets": dna.offsets,
           "protein_residues": prot.residues, "protein_offsets": prot.offsets}
    names = []
    for name, alpha, mat, o, e, glob, program, nucleic, dist_fn, double in CONFIGS:
        seqs = dna if alpha == "dna" else prot
        M = ca.load_scoring_matrix(mat)
        d, st = compute(seqs, M, o, e, glob, program, nucleic, dist_fn, double)
        out[name + "_dist"] = d
        out[name + "_stats"] = st
        out[name + "_config"] = np.array([o, e, glob, program, nucleic, dist_fn, double], dtype=np.int32)
        out[name + "_matrix"] = np.frombuffer(mat.encode(), dtype=np.uint8)
        out[name + "_alphabet"] = np.frombuffer(alpha.encode(), dtype=np.uint8)
        names.append(name)
    out["names"] = np.frombuffer(",".join(names).encode(), dtype=np.uint8)
    path = os.path.join(HERE, "pairs_v1.npz")
    np.savez_compressed(path, **out)
    print("wrote %s (%d configurations, %d DNA and %d protein sequences)" % (path, len(names), len(dna), len(prot)))


if __name__ == "__main__":
    main()
